"""Batched form (tgp_partition_batch): every graph's result equals its solo run.

The batch mixes R-MAT and Erdos-Renyi graphs of uneven size with per-graph k,
as the batched workload does, and runs them several at a time on separate
streams; each partition, part count and cut must equal what tgp_partition
returns for that graph alone, and a few are also checked against the oracle.
"""
import os

import numpy as np
import pytest

import thegraphpartitioning_amd as tgp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def batch_graphs(count, seed, nmax):
    rng = np.random.default_rng(seed)
    gs, ks = [], []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(1000), np.log(nmax)))))
        g = tgp.gen_rmat_n(n, 8, seed * 1000 + i) if i % 2 == 0 else \
            tgp.gen_er(n, 10.0, seed * 1000 + i)
        gs.append(g)
        ks.append(int(rng.choice([2, 4, 8, 16])))
    return gs, ks


@pytest.mark.parametrize("engine", ["1", "0"])
@pytest.mark.parametrize("workers", [1, 6])
def test_batch_equals_solo(workers, engine, monkeypatch):
    """engine "1": the persistent engine takes every graph; "0": per-stream workers."""
    monkeypatch.setenv("TGP_BATCH_ENGINE", engine)
    gs, ks = batch_graphs(20, 11, 30000)
    out = tgp.partition_batch(gs, ks, init=tgp.INIT_BLOCK, repair=True, workers=workers)
    for i, (g, k) in enumerate(zip(gs, ks)):
        part, pc, cut, _ = tgp.partition(g, k, init=tgp.INIT_BLOCK, repair=True)
        bp, bpc, bcut, _ = out[i]
        assert bcut == cut, f"graph {i} (n={g.n}, k={k})"
        np.testing.assert_array_equal(bp, part, err_msg=f"graph {i}")
        np.testing.assert_array_equal(bpc, pc, err_msg=f"graph {i}")


@pytest.mark.parametrize("engine,setting", [("1", ""), ("1", "2,48,2048"), ("1", "8,1536,8192"),
                                            ("0", "")])
def test_batch_at_balance_bound_equals_solo(engine, setting, monkeypatch):
    """balance_eps in the batch: engine graphs get the balance stage in the
    engine's own balance launch (at each lane width), worker graphs inside
    their solo run; both must equal the solo partition with the same bound."""
    monkeypatch.setenv("TGP_BATCH_ENGINE", engine)
    if setting:
        monkeypatch.setenv("TGP_BATCH_SETTING", setting)
    gs, ks = batch_graphs(12, 14, 20000)
    eps = 0.03
    out = tgp.partition_batch(gs, ks, init=tgp.INIT_BLOCK, repair=True, workers=4,
                              balance_eps=eps)
    monkeypatch.delenv("TGP_BATCH_SETTING", raising=False)
    for i, (g, k) in enumerate(zip(gs, ks)):
        part, pc, cut, _ = tgp.partition(g, k, init=tgp.INIT_BLOCK, repair=True,
                                         balance_eps=eps)
        bp, bpc, bcut, _ = out[i]
        assert bcut == cut, f"graph {i} (n={g.n}, k={k})"
        np.testing.assert_array_equal(bp, part, err_msg=f"graph {i}")
        np.testing.assert_array_equal(bpc, pc, err_msg=f"graph {i}")
        cap = int((1.0 + float(np.float32(eps))) * ((g.n + k - 1) // k))
        assert bpc.max() <= cap


# lanes per light vertex, heavy-row threshold, sort tile: the tuner's range
# and its ends (threshold 0: every row with an arc goes to a whole wave;
# 10^6: none does)
SETTINGS = [(2, 48, 2048), (8, 384, 8192), (4, 0, 4096), (2, 1000000, 8192), (8, 96, 2048)]


@pytest.mark.parametrize("setting", SETTINGS, ids=["-".join(map(str, x)) for x in SETTINGS])
def test_batch_settings_give_solo_results(setting, monkeypatch):
    """every engine setting gives each graph its solo result, bit for bit"""
    monkeypatch.setenv("TGP_BATCH_SETTING", ",".join(map(str, setting)))
    gs, ks = batch_graphs(16, 16, 40000)
    out = tgp.partition_batch(gs, ks, init=tgp.INIT_BLOCK, repair=True, workers=2)
    monkeypatch.delenv("TGP_BATCH_SETTING")
    for i, (g, k) in enumerate(zip(gs, ks)):
        part, pc, cut, _ = tgp.partition(g, k, init=tgp.INIT_BLOCK, repair=True)
        assert out[i][2] == cut, f"graph {i} (n={g.n}, k={k})"
        np.testing.assert_array_equal(out[i][0], part, err_msg=f"graph {i}")
        np.testing.assert_array_equal(out[i][1], pc, err_msg=f"graph {i}")


def test_autotune_table_read_by_engine(tmp_path, monkeypatch):
    """the tuner measures every setting on a sample, finds the results
    identical, writes a table; the engine then runs with that table and
    still gives the solo results"""
    monkeypatch.delenv("TGP_BATCH_SETTING", raising=False)
    gs, ks = batch_graphs(24, 17, 30000)
    path = str(tmp_path / "tune.txt")
    report, res = tgp.batch_autotune(gs, ks, path=path)
    print(report, flush=True)
    assert "results identical in every setting" in report
    assert report.count(" launch ") == 21
    text = open(path).read()
    assert text in report
    monkeypatch.setenv("TGP_BATCH_TUNE", path)
    lanes, heavy, tile = tgp.batch_tuning(10.0)
    assert f"lanes {lanes}\n" in text and f"sort_tile {tile}\n" in text
    out = tgp.partition_batch(gs, ks, init=tgp.INIT_BLOCK, repair=True, workers=2)
    for i, (g, k) in enumerate(zip(gs, ks)):
        part, pc, cut, _ = tgp.partition(g, k, init=tgp.INIT_BLOCK, repair=True)
        assert res[i][3] == 0 and res[i][2] == cut, f"graph {i}"
        np.testing.assert_array_equal(res[i][0], part, err_msg=f"graph {i}")
        assert out[i][2] == cut, f"graph {i}"
        np.testing.assert_array_equal(out[i][0], part, err_msg=f"graph {i}")
        np.testing.assert_array_equal(out[i][1], pc, err_msg=f"graph {i}")


def test_malformed_table_fails_the_batch(tmp_path, monkeypatch):
    p = tmp_path / "bad.txt"
    p.write_text("lanes 5\n")
    monkeypatch.setenv("TGP_BATCH_TUNE", str(p))
    gs, ks = batch_graphs(2, 18, 3000)
    with pytest.raises(tgp.TGPError):
        tgp.partition_batch(gs, ks, init=tgp.INIT_BLOCK, repair=True, workers=1)


def test_batch_matches_oracle(oracle):
    gs, ks = batch_graphs(6, 12, 8000)
    out = tgp.partition_batch(gs, ks, init=tgp.INIT_BLOCK, repair=True, workers=3)
    for i, (g, k) in enumerate(zip(gs, ks)):
        opart, opc, ocut, _ = oracle.partition(g.xadj, g.adjncy, g.adjwgt, k, repair=1)
        assert out[i][2] == ocut, f"graph {i}"
        np.testing.assert_array_equal(out[i][0], opart)
        np.testing.assert_array_equal(out[i][1], opc)


def test_batch_reference_semantics_and_errors():
    """Reference mode (bugs included): graphs whose coarsening the reference
    rejects fail in the batch with the same code as alone, the others match."""
    gs, ks = batch_graphs(10, 13, 20000)
    items = []
    for g, k in zip(gs, ks):
        try:
            items.append(tgp.partition(g, k, init=tgp.INIT_BLOCK))
        except tgp.TGPError as e:
            items.append(e.code)
    for i in range(len(gs)):
        try:
            r = tgp.partition_batch([gs[i]], [ks[i]], init=tgp.INIT_BLOCK, workers=1)[0]
        except tgp.TGPError as e:
            assert items[i] == e.code, f"graph {i}"
            continue
        assert not isinstance(items[i], int), f"graph {i} failed alone, not in the batch"
        assert r[2] == items[i][2]
        np.testing.assert_array_equal(r[0], items[i][0])
    ok = [i for i in range(len(gs)) if not isinstance(items[i], int)]
    out = tgp.partition_batch([gs[i] for i in ok], [ks[i] for i in ok], init=tgp.INIT_BLOCK,
                              workers=4)
    for j, i in enumerate(ok):
        assert out[j][2] == items[i][2]
        np.testing.assert_array_equal(out[j][0], items[i][0])


def test_batch_device_resident():
    gs, ks = batch_graphs(8, 14, 20000)
    dgs = [tgp.DeviceGraph(g) for g in gs]
    try:
        out = tgp.partition_batch(gs, ks, repair=True, workers=4,
                                  device_graphs=[d.pointers() for d in dgs])
    finally:
        for d in dgs:
            d.free()
    for i, (g, k) in enumerate(zip(gs, ks)):
        part, pc, cut, _ = tgp.partition(g, k, init=tgp.INIT_BLOCK, repair=True)
        assert out[i][2] == cut
        np.testing.assert_array_equal(out[i][0], part)


def test_batch_mixed_scope():
    """Graphs the engine does not take -- k above 16, a graph with a missing
    reverse arc, and one with a self loop -- go to the per-stream workers in
    the same call; every result still equals its solo run."""
    gs, ks = batch_graphs(6, 15, 6000)
    ks[1] = 32
    g = gs[2]
    adj = g.adjncy.copy()
    adj[g.xadj[10]] = (adj[g.xadj[10]] + 1) % g.n if g.xadj[11] > g.xadj[10] else adj[0]
    gs[2] = tgp.Graph(g.xadj.copy(), adj, g.adjwgt.copy())
    g = gs[3]
    adj = g.adjncy.copy()
    adj[g.xadj[5]] = 5
    gs[3] = tgp.Graph(g.xadj.copy(), adj, g.adjwgt.copy())
    out = tgp.partition_batch(gs, ks, repair=True, workers=3)
    for i, (g, k) in enumerate(zip(gs, ks)):
        part, pc, cut, _ = tgp.partition(g, k, init=tgp.INIT_BLOCK, repair=True)
        assert out[i][2] == cut, f"graph {i}"
        np.testing.assert_array_equal(out[i][0], part)
        np.testing.assert_array_equal(out[i][1], pc)


def test_batch_gpmetis_in_own_directories(tmp_path, monkeypatch, stub_gpmetis):
    """init "gpmetis" with a working directory: each graph calls gpmetis on its
    own <workdir>/graph<i>/output.graph (concurrently, from the worker
    threads); every result equals the solo run through the reference seam in
    the working directory, with the same stand-in gpmetis."""
    monkeypatch.chdir(tmp_path)
    gs, ks = batch_graphs(8, 16, 20000)
    out = tgp.partition_batch(gs, ks, init=tgp.INIT_GPMETIS, repair=True, workers=4,
                              workdir=str(tmp_path / "work"))
    calls = stub_gpmetis.read_text().splitlines()
    assert sorted(calls) == sorted(f"{tmp_path}/work/graph{i}/output.graph {k}"
                                   for i, k in enumerate(ks))
    for i, (g, k) in enumerate(zip(gs, ks)):
        part, pc, cut, _ = tgp.partition(g, k, init=tgp.INIT_GPMETIS, repair=True)
        assert (tmp_path / "work" / f"graph{i}" / "output.graph").read_bytes() == \
            (tmp_path / "output.graph").read_bytes(), f"graph {i}"
        assert out[i][2] == cut, f"graph {i}"
        np.testing.assert_array_equal(out[i][0], part)
        np.testing.assert_array_equal(out[i][1], pc)


def test_batch_small_level_pool_goes_again(monkeypatch, capfd):
    """A level pool too small for every graph at once: graphs that find it full
    are partitioned in later rounds of the persistent engine (or, if one alone
    does not fit, by the per-stream path); every result still equals its solo run."""
    monkeypatch.setenv("TGP_BATCH_ENGINE", "1")
    monkeypatch.setenv("TGP_BATCH_TRACE", "1")
    monkeypatch.setenv("TGP_BATCH_POOL_BYTES", str(16 << 20))
    gs, ks = batch_graphs(20, 11, 30000)
    out = tgp.partition_batch(gs, ks, init=tgp.INIT_BLOCK, repair=True, workers=4)
    err = capfd.readouterr().err
    line = [ln for ln in err.splitlines() if ln.startswith("# batch engine:")][-1]
    print(line, flush=True)
    rounds = int(line.split(" rounds")[0].split()[-1])
    assert rounds >= 2, line
    for i, (g, k) in enumerate(zip(gs, ks)):
        part, pc, cut, _ = tgp.partition(g, k, init=tgp.INIT_BLOCK, repair=True)
        assert out[i][2] == cut, f"graph {i}"
        np.testing.assert_array_equal(out[i][0], part, err_msg=f"graph {i}")
        np.testing.assert_array_equal(out[i][1], pc, err_msg=f"graph {i}")


def test_batch_4096_graphs_sample_equals_solo():
    """The batched workload at its full size: 4096 graphs, n log-uniform in
    [1k, 200k], half R-MAT edge factor 8, half Erdos-Renyi mean degree 10, k in
    {2, 4, 8, 16}, resident in HBM, in one call; an evenly spaced sample of 64
    results equals the solo runs bit for bit."""
    import sys
    sys.path.insert(0, ROOT)
    import bench
    gs, ks = bench.batched_workload(4096, bench.SEED)
    print(f"batch 4096: {sum(g.m for g in gs)} arcs", flush=True)
    dgs = [tgp.DeviceGraph(g) for g in gs]
    try:
        ptrs = [d.pointers() for d in dgs]
        out = tgp.partition_batch(gs, ks, init=tgp.INIT_BLOCK, repair=True, workers=16,
                                  device_graphs=ptrs)
        for i in range(0, 4096, 64):
            part, pc, cut, _ = tgp.partition(gs[i], ks[i], init=tgp.INIT_BLOCK, repair=True,
                                             device_graph=ptrs[i])
            assert out[i][2] == cut, f"graph {i}"
            np.testing.assert_array_equal(out[i][0], part, err_msg=f"graph {i}")
            np.testing.assert_array_equal(out[i][1], pc, err_msg=f"graph {i}")
    finally:
        for d in dgs:
            d.free()
