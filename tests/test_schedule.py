"""The persistent coarse step's planner (CPU: host code of libemma_rt.so).

The planner gives every item (one tile of one level step) the items it waits
for.  Here the accesses the level-step kernels make are restated item by item
-- the state buffers they read and write (double-buffered per level), the
fields that are not double-buffered (n_H, sources, reionisation times,
written by FillRad and the restriction), the injected coarse leaves and the
reflux accumulators of every (coarse leaf, face) pair -- and every pair of
conflicting accesses in the sequential order of the level steps (the
reference's recursion, advanceamr.c:243-1300 and 1309-1537) must be ordered
by a chain of dependencies.  The ticket order must be a topological order.
"""
import numpy as np
import pytest

from emma_amd import rt, synth
from emma_amd import tree as T

TILE_OCTS = 32


def radstep_steps(nl, ncoarse=1):
    """level indices of the RADSTEP recursion: per step of level k, the finer
    level takes two steps first (one per coarse step on level 0)"""
    out = []

    def rec(k):
        for _ in range(1 if k == 0 else 2):
            if k + 1 < nl:
                rec(k + 1)
            out.append(k)

    for _ in range(ncoarse):
        rec(0)
    return out


def levels_of(tree):
    return [tree.level(l) for l in range(tree.lmin, tree.lmax + 1)]


def accesses(levels, steps, coupled=True):
    """(reads, writes) per item in sequential order.  Resources: ("SR", level,
    buffer, cell) N and F, ("SC", ...) x n_H, eint, T of a state buffer,
    ("N", level, cell) n_H / sources / reionisation times (one copy), ("A",
    level, leaf, face) a reflux pair's accumulators"""
    nl = len(levels)
    nt = [(L.noct + TILE_OCTS - 1) // TILE_OCTS for L in levels]
    # reflux pairs of level k: (leaf cell, face) with an oct of level k+1 behind
    pairs_of_leaf = [dict() for _ in range(nl)]
    writer_pairs = [dict() for _ in range(nl)]  # level k+1 oct -> pairs it adds into
    for k in range(1, nl):
        nei = levels[k].nei_oct
        for o in range(levels[k].noct):
            for f in range(6):
                n = int(nei[o, f])
                if n < 0 and n > T.NEI_REFLECTIVE:
                    j = -1 - n
                    p = ("A", k - 1, j, f ^ 1)
                    pairs_of_leaf[k - 1].setdefault(j, []).append(p)
                    writer_pairs[k].setdefault(o, []).append(p)
    count = [0] * nl
    reads, writes = [], []
    for k in steps:
        p = count[k] % 2
        pc = count[k - 1] % 2 if k > 0 else None
        pf = count[k + 1] % 2 if k + 1 < nl else None
        L = levels[k]
        for t in range(nt[k]):
            R, W = set(), set()
            for o in range(t * TILE_OCTS, min(L.noct, (t + 1) * TILE_OCTS)):
                for c in range(8 * o, 8 * o + 8):
                    for part in ("SR", "SC"):  # N, F / x n_H, eint, T
                        R.add((part, k, p, c))
                        W.add((part, k, 1 - p, c))
                    W.add(("N", k, c))
                    R.add(("N", k, c))
                    if coupled:  # FillRad rewrites x n_H, eint, T of the current state
                        W.add(("SC", k, p, c))
                    ch = int(L.child[c])
                    if ch >= 0:  # restriction from the children
                        for cc in range(8 * ch, 8 * ch + 8):
                            R.add(("SR", k + 1, pf, cc))
                            R.add(("SC", k + 1, pf, cc))
                            R.add(("N", k + 1, cc))
                    for q in pairs_of_leaf[k].get(c, []):  # refluxes in, then cleared
                        R.add(q)
                        W.add(q)
                for f in range(6):
                    n = int(L.nei_oct[o, f])
                    if n >= 0:  # the neighbour's N, F (and split flag, which never changes)
                        for c in range(8 * n, 8 * n + 8):
                            R.add(("SR", k, p, c))
                    elif n > T.NEI_REFLECTIVE:  # injected coarse leaf
                        R.add(("SR", k - 1, pc, -1 - n))
                for q in writer_pairs[k].get(o, []):
                    W.add(q)
            reads.append(R)
            writes.append(W)
        count[k] += 1
    return reads, writes, nt


def check_plan(tree, steps, slots, coupled=True, chunk=None, plan=None, nqueues=1):
    """every queue's tickets follow one topological order of all items, the
    agent-scope flags mark exactly the dependencies across queues, and every
    conflicting pair of accesses (by the sequential order of the level steps)
    is ordered by a chain of dependencies; items may cover several
    consecutive tiles"""
    levels = levels_of(tree)
    if plan is None:
        plan = rt.plan_schedule(tree, steps, slots, chunk=chunk, nqueues=nqueues)
    reads_t, writes_t, nt = accesses(levels, steps, coupled)
    tbase = np.concatenate([[0], np.cumsum([nt[k] for k in steps])])
    item_step, item_tile, item_ntile = plan["item_step"], plan["item_tile"], plan["item_ntile"]
    n = len(item_step)
    # items in the sequential order: by level step, then first tile
    seq_order = np.lexsort((item_tile, item_step))
    seq = np.empty(n, dtype=np.int64)
    seq[seq_order] = np.arange(n)  # ticket -> sequential id
    covered = np.zeros(len(reads_t), dtype=np.int64)
    reads, writes = [None] * n, [None] * n
    for tk in range(n):
        i, t0, m = int(item_step[tk]), int(item_tile[tk]), int(item_ntile[tk])
        assert m >= 1 and t0 + m <= nt[steps[i]]
        ids = range(int(tbase[i]) + t0, int(tbase[i]) + t0 + m)
        covered[list(ids)] += 1
        reads[seq[tk]] = set().union(*(reads_t[j] for j in ids))
        writes[seq[tk]] = set().union(*(writes_t[j] for j in ids))
    assert (covered == 1).all(), "every tile of every level step in exactly one item"
    dep_off, deps = plan["dep_off"], plan["deps"]
    qoff, rank, flags = plan["qoff"], plan["item_rank"], plan["item_flags"]
    queue = np.searchsorted(qoff, np.arange(n), side="right") - 1  # ticket -> queue
    assert qoff[0] == 0 and qoff[-1] == n and (np.diff(qoff) >= 0).all()
    assert sorted(rank.tolist()) == list(range(n))
    for q in range(len(qoff) - 1):
        assert (np.diff(rank[qoff[q]:qoff[q + 1]]) > 0).all(), "a queue out of the start order"
    cross_out = np.zeros(n, dtype=bool)
    # ancestors as bitsets over sequential ids (items visited in rank order)
    anc = [0] * n
    for tk in np.argsort(rank):
        x = int(seq[tk])
        a = 0
        cross_in = False
        for d in deps[dep_off[tk]:dep_off[tk + 1]]:
            assert rank[d] < rank[tk], "a ticket waits for a later one of the start order"
            y = int(seq[d])
            assert y < x, "a dependency on a later item of the sequential order"
            a |= anc[y] | (1 << y)
            if queue[d] != queue[tk]:
                cross_in = True
                cross_out[d] = True
        assert bool(flags[tk] & 2) == cross_in, "acquire flag"
        anc[x] = a
    assert ((flags & 4) != 0).tolist() == cross_out.tolist(), "release flag"
    # every conflicting pair ordered: for each resource, the last write before
    # an access, and for a write the reads since that last write
    last_w, reads_since = {}, {}
    bad = []
    for x in range(n):
        for r in reads[x]:
            w = last_w.get(r)
            if w is not None and w != x and not (anc[x] >> w) & 1:
                bad.append(("RAW", r, w, x))
        for r in writes[x]:
            w = last_w.get(r)
            if w is not None and w != x and not (anc[x] >> w) & 1:
                bad.append(("WAW", r, w, x))
            for y in reads_since.get(r, ()):
                if y != x and not (anc[x] >> y) & 1:
                    bad.append(("WAR", r, y, x))
        for r in reads[x]:
            reads_since.setdefault(r, set()).add(x)
        for r in writes[x]:
            last_w[r] = x
            reads_since[r] = {x} if r in reads[x] else set()
    assert not bad, bad[:10]
    return plan


@pytest.mark.parametrize("ncoarse", [1, 2])
@pytest.mark.parametrize("slots,nqueues", [(1, 1), (7, 1), (1024, 1), (8, 8), (1024, 8)])
def test_radstep_plan_orders_every_conflict(ncoarse, slots, nqueues):
    W = synth.amr_workload(4, [300, 500, 400], "SB99_equiv_mono_TH", seed=11, sigma=1.5, grid=16)
    nl = W.tree.lmax - W.tree.lmin + 1
    plan = check_plan(W.tree, radstep_steps(nl, ncoarse), slots, nqueues=nqueues)
    assert plan["est_span"] >= plan["est_chain"] - 1e-9
    assert plan["est_span"] >= plan["est_work"] / slots - 1e-9


def test_coarserad_segments():
    """COARSERAD: the finer levels' steps (finest first) form one persistent
    run, the coarse level's step another (the transfers sit between them)"""
    W = synth.amr_workload(3, [60, 100, 80], "Starburst99_1grpT_3grpS", seed=23, sigma=1.5, grid=16)
    nl = W.tree.lmax - W.tree.lmin + 1
    for nq in (1, 8):
        check_plan(W.tree, list(range(nl - 1, 0, -1)), 64, nqueues=nq)
        check_plan(W.tree, [0], 64, nqueues=nq)


@pytest.mark.parametrize("reflect", [(), (0, 2, 4)])
def test_domain_faces(reflect):
    W = synth.amr_workload(3, [60, 100, 80], "SB99_equiv_mono_TH", seed=20, sigma=1.5, grid=16)
    tb = T.with_boundaries(W.tree, trans=range(6), reflect=reflect)
    check_plan(tb, radstep_steps(tb.lmax - tb.lmin + 1, 2), 32, nqueues=8)


def test_uniform_level_chain():
    """one complete level: step s+1 of a tile waits for its own and its
    neighbours' step s, nothing else"""
    tree = T.uniform_tree(4)
    plan = check_plan(tree, [0, 0, 0], 16)
    nt = tree.level(4).noct // TILE_OCTS
    assert len(plan["item_step"]) == 3 * nt
    first = plan["dep_off"][1:] - plan["dep_off"][:-1]
    assert (first[plan["item_step"] == 0] == 0).all()
    assert (first[plan["item_step"] > 0] > 1).all()


def test_critical_chain_goes_first():
    """a deep, narrow refinement: the fine levels' chain is the longest, so
    its items take the early tickets and the raised priority, and the list
    schedule overlaps the coarse level's independent tiles with it"""
    W = synth.amr_workload(4, [40, 60, 80], "SB99_equiv_mono_TH", seed=5, sigma=2.0, grid=16)
    nl = W.tree.lmax - W.tree.lmin + 1
    steps = radstep_steps(nl, 1)
    plan = check_plan(W.tree, steps, 64)
    step_level = np.asarray(steps)[plan["item_step"]]
    assert step_level[0] == nl - 1
    assert plan["item_prio"][step_level == nl - 1].all()
    # with enough slots the makespan is the chain, not the sum of level steps
    assert plan["est_span"] <= plan["est_chain"] + 1e-9


def test_bad_input():
    tree = T.uniform_tree(3)
    with pytest.raises(rt.EmmaRTError):
        rt.plan_schedule(tree, [1], 8)  # no such level
    with pytest.raises(rt.EmmaRTError):
        rt.plan_schedule(tree, [0], 0)  # no work-groups


@pytest.mark.parametrize("chunk", [[2, 3, 1, 1], [4, 1, 2, 5]])
def test_chunked_items(chunk):
    """items of several consecutive tiles (the rule gives them to the largest
    levels only; forced here on every level)"""
    W = synth.amr_workload(4, [300, 500, 400], "SB99_equiv_mono_TH", seed=11, sigma=1.5, grid=16)
    nl = W.tree.lmax - W.tree.lmin + 1
    plan = check_plan(W.tree, radstep_steps(nl, 2), 16, chunk=chunk, nqueues=8)
    assert plan["item_ntile"].max() == max(chunk)


def test_chunk_rule():
    """one tile per item until a level has 16 x slots tiles"""
    tree = T.uniform_tree(5)  # 16^3 = 4096 octs = 128 tiles
    for slots, expect in ((16, 1), (8, 2), (4, 4), (1, 16)):
        plan = rt.plan_schedule(tree, [0], slots)
        assert plan["item_ntile"].max() == expect, (slots, plan["item_ntile"].max())


@pytest.mark.parametrize("drop", ["same", "up", "down"])
def test_checker_catches_missing_edges(drop):
    """the checker has teeth: a plan without one class of dependency edges
    (same-level neighbours, coarser level, finer level) is rejected"""
    W = synth.amr_workload(3, [60, 100, 80], "SB99_equiv_mono_TH", seed=11, sigma=1.5, grid=16)
    nl = W.tree.lmax - W.tree.lmin + 1
    steps = radstep_steps(nl, 2)
    plan = rt.plan_schedule(W.tree, steps, 8, nqueues=8)
    lev_of_item = np.asarray(steps)[plan["item_step"]]
    off, deps = plan["dep_off"], plan["deps"]
    keep_off, keep = [0], []
    for t in range(len(plan["item_step"])):
        for d in deps[off[t]:off[t + 1]]:
            rel = lev_of_item[d] - lev_of_item[t]
            kind = "same" if rel == 0 else ("up" if rel < 0 else "down")
            if kind != drop:
                keep.append(d)
        keep_off.append(len(keep))
    # the agent-scope flags made to agree with the remaining edges, so that
    # what fails is the ordering of the accesses
    queue = np.searchsorted(plan["qoff"], np.arange(len(keep_off) - 1), side="right") - 1
    flags = plan["item_flags"] & 1
    for t in range(len(keep_off) - 1):
        for d in keep[keep_off[t]:keep_off[t + 1]]:
            if queue[d] != queue[t]:
                flags[t] |= 2
                flags[d] |= 4
    bad = dict(plan, dep_off=np.asarray(keep_off), deps=np.asarray(keep, dtype=np.int32), item_flags=flags)
    with pytest.raises(AssertionError, match="RAW|WAR|WAW"):
        check_plan(W.tree, steps, 8, plan=bad)
