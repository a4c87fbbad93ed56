"""The balance stage (tgp_options.balance_eps; no counterpart in the
reference) against a NumPy restatement of its rule, bit for bit.

The stage runs at the finest level after the reference's refinement: balance
rounds until every part holds at most cap = floor((1 + eps) * ceil(n / k))
vertices, then refine passes of positive exact integer gains within cap (see
include/tgp.h).  The restatement below starts from the plain partition (the
same pipeline without the stage) and must reach the same partition, counts,
cut and BALANCE trace record.  At full size (R-MAT 24, k = 64) the checks are
size-independent: every part within cap, the cut equal to a recount, repeat
runs identical.
"""
import math

import numpy as np
import pytest

import graphs
import thegraphpartitioning_amd as tgp

pytestmark = pytest.mark.gpu


def tallies(x, a, w, part, k):
    n = len(x) - 1
    src = np.repeat(np.arange(n), np.diff(x))
    ok = (a != src) & (a >= 0) & (a < n)
    pu = part[a[ok]]
    ok2 = (pu >= 0) & (pu < k)
    T = np.zeros((n, k), np.int64)
    np.add.at(T, (src[ok][ok2], pu[ok2]), w[ok][ok2].astype(np.int64))
    return T


def best_target(T, part, room, mode):
    """mode 0: vertices of parts over cap, any gain; mode 1: every vertex,
    gain > 0.  Returns cand, target, gain (ties: lowest part)."""
    n, k = T.shape
    own = T[np.arange(n), part]
    G = T - own[:, None]
    allowed = np.broadcast_to(room > 0, (n, k)).copy()
    allowed[np.arange(n), part] = False
    Gm = np.where(allowed, G, np.iinfo(np.int64).min)
    tgt = np.argmax(Gm, axis=1)                     # first maximum = lowest part
    gain = Gm[np.arange(n), tgt]
    has = allowed.any(axis=1)
    cand = has & ((room[part] < 0) if mode == 0 else (gain > 0))
    return cand, np.where(cand, tgt, -1), np.where(cand, gain, 0)


def rank_keep(flag, keypart, gain, limit):
    """flagged vertices ranked by (keypart, gain desc, v asc); kept while the
    rank within keypart is below limit[keypart]"""
    v = np.nonzero(flag)[0]
    order = np.lexsort((v, -gain[v], keypart[v]))
    vs = v[order]
    kp = keypart[vs]
    first = np.r_[0, np.nonzero(np.diff(kp))[0] + 1]
    start = np.repeat(first, np.diff(np.r_[first, len(vs)]))
    rank = np.arange(len(vs)) - start
    keep = np.zeros(len(flag), bool)
    keep[vs[rank < limit[kp]]] = True
    return keep


def balance_np(x, a, w, part, k, eps):
    part = part.astype(np.int64).copy()
    n = len(part)
    cap = math.floor((1.0 + float(np.float32(eps))) * ((n + k - 1) // k))
    src = np.repeat(np.arange(n), np.diff(x))
    rounds = bal = ref = 0
    while rounds < 1000:
        pc = np.bincount(part, minlength=k)
        if not (pc > cap).any():
            break
        room, excess = cap - pc, np.maximum(pc - cap, 0)
        T = tallies(x, a, w, part, k)
        cand, tgt, gain = best_target(T, part, room, 0)
        keep = rank_keep(cand, tgt, gain, room)
        keep = rank_keep(keep, part, gain, excess)
        part[keep] = tgt[keep]
        rounds += 1
        bal += int(keep.sum())
        if not keep.any():
            break
    for _ in range(30):
        pc = np.bincount(part, minlength=k)
        room = cap - pc
        T = tallies(x, a, w, part, k)
        cand, tgt, gain = best_target(T, part, room, 1)
        ok = (a != src) & cand[src] & cand[a]
        beats = (gain[a] > gain[src]) | ((gain[a] == gain[src]) & (a < src))
        lose = np.zeros(n, bool)
        lose[src[ok & beats]] = True
        win = cand & ~lose
        keep = rank_keep(win, tgt, gain, room)
        part[keep] = tgt[keep]
        ref += int(keep.sum())
        if not keep.any():
            break
    return part.astype(np.int32), rounds, bal, ref, cap


CASES = [("rmat", 12, 8, 16), ("rmat", 13, 16, 8), ("er", 3000, 10, 5), ("grid", 12, 0, 8),
         ("rmatw", 11, 8, 32)]


@pytest.mark.parametrize("spec", CASES, ids=[f"{c[0]}{c[1]}-k{c[3]}" for c in CASES])
def test_balance_stage_matches_restatement(oracle, spec):
    kind, a1, a2, k = spec
    if kind == "rmat":
        g = tgp.gen_rmat(a1, a2, 0x7C0DE5EED + a1)
    elif kind == "rmatw":
        g = tgp.gen_rmat(a1, a2, 0x7C0DE5EED + a1, "hash16")
    elif kind == "er":
        g = tgp.gen_er(a1, a2, 7)
    else:
        g = tgp.gen_grid3d(a1)
    eps = 0.03
    kw = dict(init=tgp.INIT_BLOCK, repair=True, coarsen_floor=100)
    plain = tgp.partition(g, k, **kw)
    bal = tgp.partition(g, k, balance_eps=eps, trace_cap=1 << 14, **kw)
    part, rounds, nb, nr, cap = balance_np(g.xadj, g.adjncy, g.adjwgt, plain[0], k, eps)
    np.testing.assert_array_equal(bal[0], part)
    counts = np.bincount(part, minlength=k)
    np.testing.assert_array_equal(bal[1], counts)
    assert counts.max() <= cap
    cut = oracle.edge_cut(g.xadj, g.adjncy, g.adjwgt, g.m, part, k)
    assert bal[2] == cut
    rec = [t for t in bal[3].trace if t[0] == 6]
    assert rec and rec[0][3:8] == (rounds, nb, nr, cut, int(counts.max()))


@pytest.mark.parametrize("spec", CASES[:3], ids=[f"{c[0]}{c[1]}-k{c[3]}" for c in CASES[:3]])
def test_balance_entry_equals_stage_in_partition(oracle, spec):
    """tgp_balance on the plain partition equals tgp_partition with the same
    balance_eps: partition, counts, cut and the BALANCE trace record."""
    kind, a1, a2, k = spec
    g = tgp.gen_rmat(a1, a2, 0x7C0DE5EED + a1) if kind == "rmat" else tgp.gen_er(a1, a2, 7)
    eps = 0.03
    kw = dict(init=tgp.INIT_BLOCK, repair=True, coarsen_floor=100)
    plain = tgp.partition(g, k, **kw)
    full = tgp.partition(g, k, balance_eps=eps, trace_cap=1 << 14, **kw)
    part, pc, cut, st = tgp.balance(g, plain[0], k, eps, trace_cap=64)
    np.testing.assert_array_equal(part, full[0])
    np.testing.assert_array_equal(pc, full[1])
    assert cut == full[2]
    rec = [t for t in st.trace if t[0] == 6]
    frec = [t for t in full[3].trace if t[0] == 6]
    assert rec and rec[0][3:] == frec[0][3:]
    # a second call starts balanced (no balance round) and its refine passes
    # move non-adjacent vertices by positive exact gains only: the cut cannot rise
    p2, pc2, cut2, st2 = tgp.balance(g, part, k, eps, trace_cap=64)
    rec2 = [t for t in st2.trace if t[0] == 6]
    assert rec2[0][3] == 0 and rec2[0][4] == 0
    assert cut2 <= cut and pc2.max() <= max(int(pc.max()), rec2[0][7])
    assert cut2 == oracle.edge_cut(g.xadj, g.adjncy, g.adjwgt, g.m, p2, k)
    # device-resident graph
    dg = tgp.DeviceGraph(g)
    try:
        p3, _, cut3, _ = tgp.balance(g, plain[0], k, eps, device_graph=dg.pointers())
    finally:
        dg.free()
    np.testing.assert_array_equal(p3, part)
    assert cut3 == cut


def test_balance_entry_rejects_bad_arguments():
    g = tgp.gen_er(500, 6.0, 3)
    part = np.zeros(g.n, np.int32)
    with pytest.raises(tgp.TGPError, match="balance_eps > 0"):
        tgp.balance(g, part, 4, 0.0)
    with pytest.raises(tgp.TGPError, match="k in"):
        tgp.balance(g, part, 2000, 0.03)
    # out-of-range part values count as part 0 (as the stage's recount does)
    bad = np.full(g.n, 7, np.int32)
    ok = np.zeros(g.n, np.int32)
    a = tgp.balance(g, bad, 4, 0.03)
    b = tgp.balance(g, ok, 4, 0.03)
    np.testing.assert_array_equal(a[0], b[0])


def test_balance_at_evaluation_size():
    g = tgp.gen_rmat(24, 16, 0x7C0DE5EED + 1)
    k, eps = 64, 0.03
    print(f"balance R-MAT 24: n={g.n} m={g.m} k={k}", flush=True)
    dg = tgp.DeviceGraph(g)
    try:
        kw = dict(init=tgp.INIT_BLOCK, repair=True, device_graph=dg.pointers(), balance_eps=eps)
        r1 = tgp.partition(g, k, **kw)
        r2 = tgp.partition(g, k, **kw)
    finally:
        dg.free()
    np.testing.assert_array_equal(r1[0], r2[0])
    assert r1[2] == r2[2]
    cap = math.floor((1.0 + float(np.float32(eps))) * ((g.n + k - 1) // k))
    counts = np.bincount(r1[0], minlength=k)
    assert counts.max() <= cap and counts.sum() == g.n
    np.testing.assert_array_equal(r1[1], counts)
    assert r1[2] == graphs.cut_recount(g, r1[0])
