// rt_sched.h -- the host-side planner of the persistent coarse step.
//
// A coarse step is a sequence of level steps (the RADSTEP recursion of
// advanceamr.c:243-1300, or one step per level in COARSERAD mode).  The
// persistent kernel runs it as items: one tile (32 octs) of one level step.
// The planner gives every item the earlier items it must wait for and orders
// the items into the ticket sequence the device work queue hands out.
//
// Dependencies of item (step i of level k, tile t), each on the most recent
// earlier step of the named level (none when that level has not stepped yet
// in the launch; the launch boundary orders everything before it):
//  * level k: the tiles holding t's octs and their same-level face
//    neighbours.  Their previous step wrote the state t reads (t's cells and
//    the neighbours' N, F), and read the buffer t is about to overwrite.
//  * level k-1: the tiles holding the parent cells of t's octs and the coarse
//    leaves injected across t's coarse faces.  t reads their state
//    (injection), adds refluxes into those leaves' accumulators, which their
//    step consumed and cleared, and overwrites cells their restriction read.
//  * level k+1: the tiles holding the children of t's split cells and the
//    finer octs across t's leaves' faces.  t restricts from the children and
//    takes the finer octs' refluxes.
// Every conflicting pair of accesses in the sequential order is ordered by a
// chain of these edges (the CPU test checks it against the accesses the
// kernels make).
//
// Order: a list schedule over `slots` work-groups (split evenly over the
// queues) with unit-free tile costs per level, ready items taken
// longest-remaining-chain first; each queue's tickets follow the simulated
// start times, so every queue's order agrees with one topological order of
// all items.  Items whose chain reaches the simulated end get a raised wave
// priority; items with a dependency or a dependent in another queue are
// flagged for agent-scope acquire / release.
#pragma once

#include <vector>

namespace emma {

struct SchedLevelTopo {
  int noct = 0;
  const int* nei = nullptr;    // [noct][6]: >= 0 same-level oct, < 0 coarse leaf -1-nei or a domain face
  const int* child = nullptr;  // [noct][8]: >= 0 child oct on the next level, < 0 leaf
};

// item flags: raised wave priority; the item reads results of another
// queue's items (acquire at agent scope); another queue's items read its
// results (release at agent scope)
constexpr unsigned char kSchedPrio = 1;
constexpr unsigned char kSchedAcquireAgent = 2;
constexpr unsigned char kSchedReleaseAgent = 4;

struct SchedPlan {
  int nitems = 0, nq = 1;
  std::vector<int> qoff;                              // queue q's tickets: qoff[q] .. qoff[q+1]-1
  std::vector<int> item_step, item_tile, item_ntile;  // by ticket: level step, first tile, tiles
  std::vector<unsigned char> item_prio;               // by ticket: kSched* flags
  std::vector<int> item_rank;                         // by ticket: place in the simulated start order
  std::vector<int> dep_off, deps;                     // by ticket; deps are tickets
  double est_span = 0., est_chain = 0., est_work = 0.;  // simulated makespan, longest chain, total cost
};

constexpr int kSchedTileOcts = 32;

// tiles per item of a level step with ntiles tiles: one on levels of fewer
// than 16 x slots tiles, else enough to give the step about 8 x slots items
// (at most 64), so that the per-item synchronisation stays small against
// the work on the largest levels (measured on the 256^3 level: 2, 4 and 8
// tiles per item alike, 32 slower, one tile per item slower still)
int chunk_tiles(int ntiles, int slots);

// steps[i]: level index (0 = coarsest) of the i-th level step, in the
// sequential order; tile_cost[k]: relative cost of one tile step of level k
// (<= 0: 1); slots: work-groups of the persistent grid; chunk[k] (optional,
// > 0): tiles per item of level k instead of chunk_tiles; nq: queues (one
// per XCD, each served by its XCD's work-groups, so that dependencies within
// a queue stay within one L2): a level's tiles split into nq contiguous
// ranges, except that a level with no more tiles than four times one
// queue's work-groups follows its parents' queues.  Returns 0, or -1 on bad input.
int plan_schedule(const std::vector<SchedLevelTopo>& lev, const std::vector<int>& steps,
                  const std::vector<double>& tile_cost, int slots, SchedPlan& out,
                  const std::vector<int>* chunk = nullptr, int nq = 1);

}  // namespace emma
