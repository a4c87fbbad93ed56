"""Learned move proposer: its kernels against a plain PyTorch FP32 reference of
the same operations (on BF16-rounded inputs, since the kernels take BF16), the
proposal rule against a NumPy restatement, and the learned refinement's output
against the size-independent checks (parts in range, counts summing to n, the
reported cut equal to a recount, repeat runs identical).

Tolerances: products of BF16 values are exact in FP32, so GEMM results differ
from the reference only by FP32 summation order (rtol 1e-4 of the row scale);
outputs written back as BF16 carry BF16 rounding (rtol 2^-8).
"""
import ctypes as C

import numpy as np
import pytest

import thegraphpartitioning_amd as tgp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
learn = pytest.importorskip("thegraphpartitioning_amd.learn")


def bf16_round(x):
    return torch.as_tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float32)


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("shape", [(100, 70, 45, 0, 0), (64, 64, 64, 1, 0), (37, 130, 200, 0, 1),
                                   (300, 16, 96, 1, 1)])
def test_gemm_matches_fp32(shape):
    M, N, K, ta, tb = shape
    g = torch.Generator().manual_seed(M * N + K)
    A = bf16_round(torch.randn(K, M, generator=g) if ta else torch.randn(M, K, generator=g))
    B = bf16_round(torch.randn(N, K, generator=g) if tb else torch.randn(K, N, generator=g))
    Cin = torch.randn(M, N, generator=g)
    bias = torch.randn(N, generator=g)
    ref = (A.t() if ta else A) @ (B.t() if tb else B)
    ref = torch.relu(0.5 * ref + 2.0 * Cin + bias)
    dA = A.to(torch.bfloat16).cuda()
    dB = B.to(torch.bfloat16).cuda()
    dC = Cin.clone().cuda()
    dCb = torch.empty(M, N, dtype=torch.bfloat16, device="cuda")
    L = learn._lib()
    rc = L.tgpl_gemm(learn._p(dA), dA.shape[1], ta, learn._p(dB), dB.shape[1], tb, learn._p(dC), N,
                     learn._p(dCb), M, N, K, 0.5, 2.0, learn._p(bias.cuda()), 1, stream())
    assert rc == 0
    torch.cuda.synchronize()
    got = dC.cpu()
    scale = ref.abs().max().item() + 1.0
    assert (got - ref).abs().max().item() <= 1e-4 * scale * K ** 0.5
    assert torch.allclose(dCb.cpu().float(), got, rtol=2 ** -8, atol=1e-6)


@pytest.mark.parametrize("shape", [(64, 64, 20000, 1, 0), (150, 40, 70001, 1, 0),
                                   (130, 64, 9000, 1, 1)])
def test_split_k_gemm(shape):
    """Weight-gradient shapes (few output tiles, long K) through the split-K
    path: against float64, each entry within K * 2^-23 * (|A| |B|) of the
    exact product (FP32 sums of exact BF16 products in any order), and equal
    on a repeat (the slices are added in a fixed order)."""
    M, N, K, ta, tb = shape
    g = torch.Generator().manual_seed(M + N + K)
    A = bf16_round(torch.randn(K, M, generator=g) if ta else torch.randn(M, K, generator=g))
    B = bf16_round(torch.randn(N, K, generator=g) if tb else torch.randn(K, N, generator=g))
    opA = (A.t() if ta else A).double()
    opB = (B.t() if tb else B).double()
    ref = (opA @ opB).numpy()
    bound = K * 2.0 ** -23 * (opA.abs() @ opB.abs()).numpy() + 1e-30
    dA = A.to(torch.bfloat16).cuda()
    dB = B.to(torch.bfloat16).cuda()
    work = torch.empty(4 << 20, device="cuda")
    L = learn._lib()
    outs = []
    for _ in range(2):
        dC = torch.empty(M, N, device="cuda")
        assert L.tgpl_gemm_ws(learn._p(dA), dA.shape[1], ta, learn._p(dB), dB.shape[1], tb,
                              learn._p(dC), N, None, M, N, K, 1.0, 0.0, None, 0,
                              learn._p(work), work.numel(), stream()) == 0
        torch.cuda.synchronize()
        outs.append(dC.cpu().double().numpy())
    assert (np.abs(outs[0] - ref) <= bound).all(), np.max(np.abs(outs[0] - ref) / bound)
    np.testing.assert_array_equal(outs[0], outs[1])


def small_graph():
    return tgp.gen_rmat(9, 8, 5, "hash16")


def hub_graph():
    """Two hub rows of 4999 and 4998 arcs: past the 2048-arc limit above which
    a row is aggregated by a block of 16 waves instead of one wave."""
    import graphs
    xadj, adj, w = graphs.hub_graph(5000, 2, 6000, 3, wmax=5)
    return tgp.Graph(xadj, adj, w)


@pytest.mark.parametrize("which", ["rmat9", "hubs"])
def test_mean_aggregation_and_adjoint(which):
    """Against float64 sparse products.  Tolerance: an FP32 sum of d terms in
    any order is within d * 2^-23 * sum|terms| of the exact sum (plus the
    final division's half ulp), so each entry is held to that bound."""
    import scipy.sparse as sp
    g = small_graph() if which == "rmat9" else hub_graph()
    dg = learn.DevGraph(g)
    F = 64 if which == "hubs" else 48
    X = bf16_round(torch.randn(g.n, F, generator=torch.Generator().manual_seed(1)))
    deg = np.diff(g.xadj)
    src = np.repeat(np.arange(g.n), deg)
    A = sp.csr_matrix((np.ones(len(src)), (src, g.adjncy.astype(np.int64))), shape=(g.n, g.n))
    D = np.maximum(deg, 1).astype(np.float64)
    Xd = X.double().numpy()
    fwd = (A @ Xd) / D[:, None]
    fwd_abs = (A @ np.abs(Xd)) / D[:, None]
    adj = A.T @ (Xd / D[:, None])
    adj_abs = A.T @ (np.abs(Xd) / D[:, None])
    cnt_in = np.asarray(A.sum(0)).ravel()
    L = learn._lib()
    for transpose, ref, mag, terms in ((0, fwd, fwd_abs, deg), (1, adj, adj_abs, cnt_in)):
        out = torch.empty(g.n, F, dtype=torch.bfloat16, device="cuda")
        outf = torch.empty(g.n, F, device="cuda")
        assert L.tgpl_mean_agg(g.n, learn._p(dg.xadj), learn._p(dg.adj),
                               learn._p(X.to(torch.bfloat16).cuda()), F, transpose,
                               learn._p(out), learn._p(outf), stream()) == 0
        torch.cuda.synchronize()
        got = outf.cpu().double().numpy()
        bound = (np.maximum(terms, 1)[:, None] + 2) * 2.0 ** -23 * mag + 1e-30
        assert (np.abs(got - ref) <= bound).all(), np.max(np.abs(got - ref) / bound)
        np.testing.assert_array_equal(out.cpu().float().numpy(),
                                      bf16_round(torch.from_numpy(got).float()).numpy())


@pytest.mark.parametrize("k,which", [(8, "rmat9"), (130, "rmat9"), (16, "hubs")])
def test_features_and_proposals(k, which):
    g = small_graph() if which == "rmat9" else hub_graph()
    dg = learn.DevGraph(g)
    part = (np.arange(g.n) * k // g.n).astype(np.int32)
    F = learn.feature_count(k)
    X = torch.empty(g.n, F, dtype=torch.bfloat16, device="cuda")
    tally = torch.empty(g.n, k, dtype=torch.int32, device="cuda")
    dpart = torch.from_numpy(part).cuda()
    L = learn._lib()
    assert L.tgpl_features(g.n, learn._p(dg.xadj), learn._p(dg.adj), learn._p(dg.w),
                           learn._p(dpart), k, F, learn._p(X), learn._p(tally), stream()) == 0
    torch.cuda.synchronize()
    src = np.repeat(np.arange(g.n), np.diff(g.xadj))
    T = np.zeros((g.n, k), np.int64)
    np.add.at(T, (src, part[g.adjncy]), g.adjwgt)
    np.testing.assert_array_equal(tally.cpu().numpy(), T)
    deg = T.sum(1)
    x = X.cpu().float().numpy()
    np.testing.assert_allclose(x[:, :k], bf16_round(T / np.maximum(deg, 1)[:, None]).numpy())
    np.testing.assert_array_equal(x[:, k:2 * k], np.eye(k)[part])
    assert (x[:, 2 * k + 1] == 1).all() and (x[:, 2 * k + 2:] == 0).all()
    # proposals: best other part by logit (lowest on ties), admitted only with a
    # positive exact gain and a logit within the margin of the own part's
    Z = torch.randn(g.n, k, generator=torch.Generator().manual_seed(4))
    Z[:5] = 0.0
    gain = torch.empty(g.n, device="cuda")
    tgt = torch.empty(g.n, dtype=torch.int32, device="cuda")
    margin = 0.5
    assert L.tgpl_propose(g.n, k, learn._p(Z.cuda()), learn._p(dpart), learn._p(tally), margin,
                          learn._p(gain), learn._p(tgt), stream()) == 0
    torch.cuda.synchronize()
    Zn = Z.numpy()
    tgt_h = tgt.cpu().numpy()
    gain_h = gain.cpu().numpy()
    for v in range(g.n):
        cur = part[v]
        order = [q for q in range(k) if q != cur]
        best = max(order, key=lambda q: (Zn[v, q], -q))
        d = T[v, best] - T[v, cur]
        ok = Zn[v, best] > Zn[v, cur] - np.float32(margin) and d > 0
        assert tgt_h[v] == best
        assert (gain_h[v] == d) if ok else np.isnan(gain_h[v])


def reference_forward(model, g, part):
    """The same GNN in PyTorch FP32 on CPU, BF16 rounding where the kernels
    store BF16 (features, aggregates, hidden activations, weights)."""
    k, F = model.k, model.F
    src = np.repeat(np.arange(g.n), np.diff(g.xadj))
    T = np.zeros((g.n, k))
    np.add.at(T, (src, part[g.adjncy]), g.adjwgt)
    deg = T.sum(1)
    X = np.zeros((g.n, F), np.float32)
    X[:, :k] = T / np.maximum(deg, 1)[:, None]
    X[:, k:2 * k] = np.eye(k)[part]
    X[:, 2 * k] = np.log2(1 + np.diff(g.xadj)) / 16
    X[:, 2 * k + 1] = 1
    X = bf16_round(X)
    A = torch.zeros(g.n, g.n)
    A[torch.from_numpy(src), torch.from_numpy(g.adjncy.astype(np.int64))] = 1.0
    D = torch.from_numpy(np.maximum(np.diff(g.xadj), 1).astype(np.float32))[:, None]
    W = {n: w.cpu() for n, w in model.W.items()}
    Wb = {n: bf16_round(w) for n, w in W.items()}
    P = bf16_round(X @ Wb["W1b"])
    H = bf16_round(torch.relu(X @ Wb["W1a"] + (A @ P) / D + W["b1"]))
    M1 = bf16_round((A @ H) / D)
    return H @ Wb["W2a"] + M1 @ Wb["W2b"] + W["b2"]


def test_forward_matches_reference():
    g = small_graph()
    k = 8
    model = learn.GNN(k, seed=3)
    part = (np.arange(g.n) * k // g.n).astype(np.int32)
    Z, _ = model.forward(learn.DevGraph(g), torch.from_numpy(part).cuda())
    ref = reference_forward(model, g, part)
    torch.cuda.synchronize()
    assert torch.allclose(Z.cpu(), ref, rtol=2e-2, atol=2e-2)


def test_training_reduces_loss_on_plain_traces():
    gs = [tgp.gen_rmat(11, 8, 30 + i) for i in range(2)]
    k = 8
    data = learn.collect_traces(gs, k, init=tgp.INIT_BLOCK, repair=True, coarsen_floor=100)
    assert all(len(p) > 0 for _, p in data)
    model = learn.GNN(k, seed=0)
    dgs = [learn.DevGraph(g) for g, _ in data]
    g0, pairs0 = data[0]
    b0 = torch.from_numpy(pairs0[0][0]).cuda()
    a0 = torch.from_numpy(pairs0[0][1]).cuda()
    before = model.evaluate(dgs[0], b0, a0)[0]
    for _ in range(15):
        for dg, (g, pairs) in zip(dgs, data):
            for b, a in pairs[:4]:
                model.train_step(dg, torch.from_numpy(b).cuda(), torch.from_numpy(a).cuda())
    after = model.evaluate(dgs[0], b0, a0)[0]
    assert after < before


def test_learned_partition_is_valid_and_deterministic():
    g = tgp.gen_rmat(12, 8, 41)
    k = 8
    data = learn.collect_traces([g], k, init=tgp.INIT_BLOCK, repair=True, coarsen_floor=100)
    model = learn.GNN(k, seed=1)
    dg = learn.DevGraph(g)
    for _ in range(10):
        for b, a in data[0][1][:6]:
            model.train_step(dg, torch.from_numpy(b).cuda(), torch.from_numpy(a).cuda())
    h = model.handle(max_passes=10, min_admit=0.0)
    kw = dict(init=tgp.INIT_BLOCK, repair=True, coarsen_floor=100, learned=h)
    r1 = tgp.partition(g, k, **kw)
    r2 = tgp.partition(g, k, **kw)
    part, pc, cut, _ = r1
    np.testing.assert_array_equal(r1[0], r2[0])
    assert r1[2] == r2[2]
    assert part.min() >= 0 and part.max() < k and pc.sum() == g.n
    src = np.repeat(np.arange(g.n), np.diff(g.xadj))
    recount = int(g.adjwgt[part[src] != part[g.adjncy]].astype(np.int64).sum()) // 2
    assert cut == recount


@pytest.mark.parametrize("which", ["rmat24", "grid256"])
def test_learned_at_evaluation_size(which):
    """The learned form on its evaluation graphs -- R-MAT scale 24 edge factor
    16 at k = 64 and the 256^3 grid at k = 128 (wide mode) -- with a model
    trained briefly on small graphs of the same family: parts in range, counts
    summing to n, the reported cut equal to a host recount, repeat runs
    identical."""
    import graphs
    if which == "rmat24":
        k, train, wide = 64, tgp.gen_rmat(14, 16, 7), False
        g = tgp.gen_rmat(24, 16, 0x7C0DE5EED + 1)
    else:
        k, train, wide = 128, tgp.gen_grid3d(24), True
        g = tgp.gen_grid3d(256)
    print(f"learned {which}: n={g.n} m={g.m} k={k}", flush=True)
    data = learn.collect_traces([train], k, init=tgp.INIT_BLOCK, repair=True, wide=wide)
    model = learn.GNN(k, seed=2)
    dtrain = learn.DevGraph(train)
    for b, a in data[0][1][:8]:
        model.train_step(dtrain, torch.from_numpy(b).cuda(), torch.from_numpy(a).cuda())
    del dtrain
    h = model.handle(max_passes=4, min_admit=0.0)
    dg = tgp.DeviceGraph(g)
    try:
        kw = dict(init=tgp.INIT_BLOCK, repair=True, wide=wide, learned=h,
                  device_graph=dg.pointers())
        r1 = tgp.partition(g, k, **kw)
        r2 = tgp.partition(g, k, **kw)
    finally:
        dg.free()
    part, pc, cut, _ = r1
    np.testing.assert_array_equal(r1[0], r2[0])
    np.testing.assert_array_equal(r1[1], r2[1])
    assert r1[2] == r2[2]
    assert part.min() >= 0 and part.max() < k and int(pc.sum()) == g.n
    assert cut == graphs.cut_recount(g, part)
