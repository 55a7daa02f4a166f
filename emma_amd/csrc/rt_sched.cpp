// rt_sched.cpp -- the planner of the persistent coarse step (see rt_sched.h).
// Host code only.
#include "rt_sched.h"

#include <algorithm>
#include <cstdint>
#include <functional>
#include <queue>
#include <utility>

namespace emma {

namespace {

constexpr int kNeiDomainMax = -2147483647;  // nei entries at or below: domain faces (transmissive / reflective)

using Sets = std::vector<std::vector<int>>;

void dedupe(Sets& s) {
  for (auto& v : s) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
}

int ntiles_of(int noct) { return (noct + kSchedTileOcts - 1) / kSchedTileOcts; }

}  // namespace

int chunk_tiles(int ntiles, int slots) {
  const int c = ntiles / (8 * std::max(slots, 1));
  return std::max(1, std::min(c, 64));
}

int plan_schedule(const std::vector<SchedLevelTopo>& lev, const std::vector<int>& steps,
                  const std::vector<double>& tile_cost, int slots, SchedPlan& out, const std::vector<int>* chunk,
                  int nq) {
  const int nl = (int)lev.size();
  out = SchedPlan();
  if (slots < 1) return -1;
  std::vector<int> nt(nl);
  for (int k = 0; k < nl; k++) {
    if (lev[k].noct < 0 || (lev[k].noct > 0 && (!lev[k].nei || !lev[k].child))) return -1;
    nt[k] = ntiles_of(lev[k].noct);
  }
  for (int k : steps)
    if (k < 0 || k >= nl || nt[k] == 0) return -1;
  // tiles per item and items per level step
  std::vector<int> cs(nl), ni(nl);
  for (int k = 0; k < nl; k++) {
    cs[k] = (chunk && k < (int)chunk->size() && (*chunk)[k] > 0) ? (*chunk)[k] : chunk_tiles(nt[k], slots);
    ni[k] = (nt[k] + cs[k] - 1) / cs[k];
  }
  // tile relations: same[k][t] (same-level neighbours and t), up[k][t] (tiles
  // of level k-1 t is coupled to), down[k][t] (tiles of level k+1 coupled to t)
  std::vector<Sets> same(nl), up(nl), down(nl);
  for (int k = 0; k < nl; k++) {
    same[k].assign(nt[k], {});
    up[k].assign(nt[k], {});
    down[k].assign(nt[k], {});
  }
  for (int k = 0; k < nl; k++) {
    const SchedLevelTopo& L = lev[k];
    for (int o = 0; o < L.noct; o++) {
      const int t = o / kSchedTileOcts;
      same[k][t].push_back(t);
      for (int f = 0; f < 6; f++) {
        const int n = L.nei[o * 6 + f];
        if (n >= 0) {
          if (n >= L.noct) return -1;
          const int tn = n / kSchedTileOcts;
          same[k][t].push_back(tn);
          same[k][tn].push_back(t);
        } else if (n > kNeiDomainMax) {  // injected coarse leaf -1-n
          if (k == 0) return -1;
          const int j = -1 - n;
          if (j >= lev[k - 1].noct * 8) return -1;
          up[k][t].push_back((j >> 3) / kSchedTileOcts);
        }
      }
    }
    if (k + 1 < nl) {  // parents of the next level's octs
      for (int c = 0; c < L.noct * 8; c++) {
        const int ch = L.child[c];
        if (ch < 0) continue;
        if (ch >= lev[k + 1].noct) return -1;
        up[k + 1][ch / kSchedTileOcts].push_back((c >> 3) / kSchedTileOcts);
      }
    }
  }
  for (int k = 0; k < nl; k++) {
    dedupe(same[k]);
    dedupe(up[k]);
    if (k > 0)
      for (int t = 0; t < nt[k]; t++)
        for (int u : up[k][t]) down[k - 1][u].push_back(t);
  }
  for (int k = 0; k < nl; k++) dedupe(down[k]);
  // the same relations between items (runs of cs[k] tiles)
  auto to_items = [&](const Sets& rel, int k, int ko) {
    Sets r(ni[k]);
    for (int t = 0; t < nt[k]; t++)
      for (int u : rel[t]) r[t / cs[k]].push_back(u / cs[ko]);
    dedupe(r);
    return r;
  };
  for (int k = 0; k < nl; k++) {
    same[k] = to_items(same[k], k, k);
    if (k > 0) up[k] = to_items(up[k], k, k - 1);
    if (k + 1 < nl) down[k] = to_items(down[k], k, k + 1);
  }

  // items in the sequential order, their dependencies (sequential ids)
  const int ns = (int)steps.size();
  std::vector<long long> base(ns + 1, 0);
  for (int i = 0; i < ns; i++) base[i + 1] = base[i] + ni[steps[i]];
  if (base[ns] > (1LL << 30)) return -1;
  const int n = (int)base[ns];
  std::vector<int> seq_step(n), seq_tile(n), seq_ntile(n);
  std::vector<double> cost(n);
  std::vector<int> doff(n + 1, 0);
  std::vector<int> dep;
  std::vector<int> last(nl, -1);  // most recent step of each level
  for (int i = 0; i < ns; i++) {
    const int k = steps[i];
    const double tc = (k < (int)tile_cost.size() && tile_cost[k] > 0.) ? tile_cost[k] : 1.;
    const int ps = last[k], pu = k > 0 ? last[k - 1] : -1, pd = k + 1 < nl ? last[k + 1] : -1;
    for (int t = 0; t < ni[k]; t++) {
      const int x = (int)base[i] + t;
      seq_step[x] = i;
      seq_tile[x] = t * cs[k];
      seq_ntile[x] = std::min(cs[k], nt[k] - t * cs[k]);
      cost[x] = tc * seq_ntile[x];
      if (ps >= 0)
        for (int u : same[k][t]) dep.push_back((int)base[ps] + u);
      if (pu >= 0)
        for (int u : up[k][t]) dep.push_back((int)base[pu] + u);
      if (pd >= 0)
        for (int u : down[k][t]) dep.push_back((int)base[pd] + u);
      doff[x + 1] = (int)dep.size();
    }
    last[k] = i;
  }
  // successors, longest remaining chain (deps point to earlier ids, so the
  // reverse sequential order is a reverse topological order)
  std::vector<int> soff(n + 1, 0), indeg(n, 0);
  for (int x = 0; x < n; x++) {
    indeg[x] = doff[x + 1] - doff[x];
    for (int e = doff[x]; e < doff[x + 1]; e++) soff[dep[e] + 1]++;
  }
  for (int x = 0; x < n; x++) soff[x + 1] += soff[x];
  std::vector<int> succ(soff[n]), fill(soff.begin(), soff.end() - 1);
  for (int x = 0; x < n; x++)
    for (int e = doff[x]; e < doff[x + 1]; e++) succ[fill[dep[e]]++] = x;
  std::vector<double> chain(n, 0.);
  double work = 0.;
  for (int x = n - 1; x >= 0; x--) {
    double m = 0.;
    for (int e = soff[x]; e < soff[x + 1]; e++) m = std::max(m, chain[succ[e]]);
    chain[x] = cost[x] + m;
    work += cost[x];
  }
  // queue of each tile: on the coarsest level and on levels with more tiles
  // than four times one queue's work-groups, nq contiguous ranges of tiles
  // (one work-group per tile and step would wait for two rounds); on smaller
  // levels the queue of the tile holding the parent of the tile's first oct
  // (so that a small level's steps and their coupling to the coarser level
  // stay in one XCD's L2).  An item takes the queue of its first tile; a tile
  // has the same queue in every step of its level.
  const int nqq = std::max(1, nq);
  const int per_q = std::max(1, slots / nqq);
  std::vector<std::vector<int>> tq(nl);
  for (int k = 0; k < nl; k++) {
    tq[k].assign(nt[k], 0);
    if (k == 0 || nt[k] > 4 * per_q) {
      for (int t = 0; t < nt[k]; t++) tq[k][t] = (int)((long long)t * nqq / nt[k]);
    } else {
      std::vector<int> parent(lev[k].noct, 0);  // parent cell's oct, level k-1
      for (int c = 0; c < lev[k - 1].noct * 8; c++)
        if (lev[k - 1].child[c] >= 0) parent[lev[k - 1].child[c]] = c >> 3;
      for (int t = 0; t < nt[k]; t++) tq[k][t] = tq[k - 1][parent[t * kSchedTileOcts] / kSchedTileOcts];
    }
  }
  std::vector<int> q_of(n);
  for (int x = 0; x < n; x++) q_of[x] = tq[steps[seq_step[x]]][seq_tile[x]];
  // list schedule, the work-groups split evenly over the queues
  using Ready = std::pair<double, int>;  // (chain, -id): longest chain first, then sequential order
  std::vector<std::priority_queue<Ready>> ready(nqq);
  using Run = std::pair<double, int>;  // (finish, id), earliest first
  std::priority_queue<Run, std::vector<Run>, std::greater<Run>> running;
  for (int x = 0; x < n; x++)
    if (indeg[x] == 0) ready[q_of[x]].push({chain[x], -x});
  std::vector<int> order;
  order.reserve(n);
  std::vector<double> start(n, 0.);
  double now = 0., span = 0.;
  std::vector<int> free_slots(nqq, std::max(1, slots / nqq));
  while ((int)order.size() < n || !running.empty()) {
    for (int q = 0; q < nqq; q++) {
      while (free_slots[q] > 0 && !ready[q].empty()) {
        const int x = -ready[q].top().second;
        ready[q].pop();
        start[x] = now;
        order.push_back(x);
        running.push({now + cost[x], x});
        free_slots[q]--;
      }
    }
    if (running.empty()) return -1;  // unreachable items (cannot happen: deps point backwards)
    const Run r = running.top();
    running.pop();
    now = r.first;
    span = std::max(span, now);
    free_slots[q_of[r.second]]++;
    for (int e = soff[r.second]; e < soff[r.second + 1]; e++)
      if (--indeg[succ[e]] == 0) ready[q_of[succ[e]]].push({chain[succ[e]], -succ[e]});
  }
  double longest = 0.;
  for (int x = 0; x < n; x++) longest = std::max(longest, chain[x]);
  // tickets: queue by queue, each in the simulated start order (so each
  // queue's order agrees with one topological order of all items)
  std::vector<int> rank(n), ticket(n), tord;
  tord.reserve(n);
  for (int i = 0; i < n; i++) rank[order[i]] = i;
  out.qoff.assign(nqq + 1, 0);
  for (int q = 0; q < nqq; q++) {
    for (int i = 0; i < n; i++)
      if (q_of[order[i]] == q) tord.push_back(order[i]);
    out.qoff[q + 1] = (int)tord.size();
  }
  for (int i = 0; i < n; i++) ticket[tord[i]] = i;
  out.nitems = n;
  out.nq = nqq;
  out.item_step.resize(n);
  out.item_tile.resize(n);
  out.item_ntile.resize(n);
  out.item_prio.resize(n);
  out.item_rank.resize(n);
  out.dep_off.assign(n + 1, 0);
  out.deps.reserve(dep.size());
  std::vector<unsigned char> cross_out(n, 0);
  for (int x = 0; x < n; x++)
    for (int e = doff[x]; e < doff[x + 1]; e++)
      if (q_of[dep[e]] != q_of[x]) cross_out[dep[e]] = 1;
  for (int i = 0; i < n; i++) {
    const int x = tord[i];
    out.item_step[i] = seq_step[x];
    out.item_tile[i] = seq_tile[x];
    out.item_ntile[i] = seq_ntile[x];
    out.item_rank[i] = rank[x];
    bool cross_in = false;
    for (int e = doff[x]; e < doff[x + 1]; e++) {
      out.deps.push_back(ticket[dep[e]]);
      cross_in = cross_in || q_of[dep[e]] != q_of[x];
    }
    out.dep_off[i + 1] = (int)out.deps.size();
    out.item_prio[i] = (unsigned char)(((start[x] + chain[x] >= 0.8 * span) ? kSchedPrio : 0) |
                                       (cross_in ? kSchedAcquireAgent : 0) | (cross_out[x] ? kSchedReleaseAgent : 0));
  }
  out.est_span = span;
  out.est_chain = longest;
  out.est_work = work;
  return 0;
}

}  // namespace emma
