"""Benchmark: end-to-end partitioning throughput on one MI355X.

Workload (one step = one full multilevel partition of the graph, the CSR
already resident in HBM): seeded R-MAT scale 24, edge factor 16, symmetrized
and deduplicated, isolated vertices removed (8,871,471 vertices, 520,761,282
arcs), unit weights, int32 CSR, k = 64.  The reference's exact coarsening
cannot finish R-MAT graphs (it stalls, one merge per level, and ends in its
"Invalid new_coarsed_edge" failure), so the step runs the opt-in repaired
coarsening; matching, refinement, move selection and cut are the
reference's.  The initial partition of the coarsest graph is the
deterministic block rule (no METIS on the machine).

Prints one JSON line: M edges/s (level-0 arcs x steps / seconds), the roofline
of the dominant kernel (of token passing, k_token, and the chained matching,
k_match_chain, the one with the larger share of the step; the other is
reported as roofline.next_kernel) from HIP-event timing of that kernel alone,
the whole step's algorithmic bytes by the per-kernel model over the step time,
and the oracle's CPU throughput on a bounded sample of the same workload family
(the same generator, edge factor, k and semantics at R-MAT scale 20 on every
host thread, and scale 18 on one).  `traffic` (HBM bytes per launch of each
kernel) comes from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed
under profiles/ (<kernel>_rmat24_traffic.json), scaled by the same counters'
reading for a plain copy of known size (`--copy-probe`).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED = 0x7C0DE5EED + 1
HBM_PEAK_GBS = 8000.0
# Single-kernel families the driver brackets with HIP events on its stream
# (family -> kernel); the roofline line is the one with the larger share of
# the profiled step, the other is reported beside it.
ROOFLINE_FAMILIES = {"token": "k_token", "match_chain": "k_match_chain"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/<kernel>_rmat24_traffic.json), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", f"{kernel}_rmat24_traffic.json")) as f:
            t = json.load(f)
        return t["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def copy_probe(mib):
    """A plain device-to-device copy of known size, for calibrating FETCH_SIZE and
    WRITE_SIZE (profile this run with --pmc and compare with mib * 1024 KiB)."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "copies": 3, "bytes_read_per_copy": n * 4,
                      "bytes_written_per_copy": n * 4}), flush=True)


def host_cpu():
    """The host's logical CPU count and model name (for the CPU baseline)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def gpmetis_baseline(g, k):
    """The reference's CPU path, gpmetis k-way with METIS defaults (ufactor 30 =
    3 %), on the METIS text of g: median wall time of 3 runs, or None when no
    gpmetis is installed (nothing else is substituted for it)."""
    import shutil
    import subprocess
    import tempfile
    import thegraphpartitioning_amd as tgp
    exe = shutil.which("gpmetis")
    if not exe:
        return None
    text, _ = tgp.metis_text(g)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "graph.metis")
        with open(path, "wb") as f:
            f.write(text)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            subprocess.run([exe, path, str(k)], check=True, stdout=subprocess.DEVNULL)
            times.append(time.perf_counter() - t0)
    return sorted(times)[1]


def cpu_baseline(scale, ef, k, scale_one):
    """The oracle end-to-end, repaired mode: on R-MAT(scale, ef) with every
    host thread (its order-independent loops use OpenMP; the matching, move
    buffering and selection stay sequential), and on the smaller
    R-MAT(scale_one, ef) with one thread; plus gpmetis when installed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    import thegraphpartitioning_amd as tgp
    nproc, model = host_cpu()
    runs = {}
    for threads, sc in ((nproc, scale), (1, scale_one)):
        g = tgp.gen_rmat(sc, ef, SEED)
        oracle_ffi.set_threads(threads)
        t0 = time.perf_counter()
        oracle_ffi.partition(g.xadj, g.adjncy, g.adjwgt, k, repair=1)
        dt = time.perf_counter() - t0
        runs[threads] = (dt, g.m / dt / 1e6, g)
        print(f"# oracle R-MAT {sc}: {dt:.1f}s with {threads} threads", file=sys.stderr,
              flush=True)
    g = runs[nproc][2]
    g1 = runs[1][2]
    gm = gpmetis_baseline(g, k)
    return {"value": round(runs[nproc][1], 4), "unit": "M edges/s", "cores": nproc,
            "kind": "port",
            "sample": f"oracle end-to-end, same repaired semantics, k={k}: R-MAT scale {scale} "
                      f"ef {ef} ({g.n} vertices, {g.m} arcs) in {runs[nproc][0]:.1f} s with "
                      f"{nproc} threads; value_1_thread on R-MAT scale {scale_one} ef {ef} "
                      f"({g1.n} vertices, {g1.m} arcs) in {runs[1][0]:.1f} s",
            "value_1_thread": round(runs[1][1], 4), "nproc": nproc, "cpu_model": model,
            "gpmetis": ("unavailable (not installed on this host)" if gm is None else
                        {"value": round(g.m / gm / 1e6, 4), "unit": "M edges/s", "cores": 1,
                         "median_seconds": round(gm, 3)})}


def model_bytes(trace, o=8):
    """Algorithmic bytes of one partition by the per-kernel model (offsets of o
    bytes, ids and weights of 4): per level built, matching o*n + 12m + 4n,
    relabel 8n + 8n_c, contraction 3o*n_c + 12m + 8m_c, projection 12n_c + 8n;
    per refinement pass 33m + (3o + 24)n + 77M + 4S (gains, token passing,
    compaction, sort, selection and commit, edge cut), with m the level's arcs
    and M, S the pass's buffered and selected moves."""
    total = 0.0
    levels = {}
    for t in trace:
        kind, level = t[0], t[1]
        if kind == 1:                                   # LEVEL: n, m, n_c, m_c
            n, m, nc, mc = t[4], t[5], t[6], t[7]
            total += o * n + 12 * m + 4 * n + 8 * n + 8 * nc + 3 * o * nc + 12 * m + 8 * mc
            total += 12 * nc + 8 * n
            levels[level] = (n, m)
            levels[level + 1] = (nc, mc)
    for t in trace:
        if t[0] == 3:                                   # PASS: M, S
            n, m = levels.get(t[1], (0, 0))
            total += 33 * m + (3 * o + 24) * n + 77 * t[4] + 4 * t[5]
    return total


def batched_workload(count, seed):
    """Independent graphs of uneven size: vertex counts log-uniform in [1k, 200k],
    even indices R-MAT edge factor 8 and odd ones Erdos-Renyi mean degree 10
    (both symmetrized, isolated vertices removed), unit weights, k drawn from
    {2, 4, 8, 16}.  Generated on host threads (the generator releases the GIL)."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    import thegraphpartitioning_amd as tgp
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(1000), np.log(200000), count)).round().astype(int)
    ks = rng.choice([2, 4, 8, 16], count).astype(int)

    def make(i):
        s = seed * 10007 + i
        return tgp.gen_rmat_n(int(ns[i]), 8, s) if i % 2 == 0 else tgp.gen_er(int(ns[i]), 10.0, s)
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        gs = list(ex.map(make, range(count)))
    return gs, [int(k) for k in ks]


def bench_batched(args):
    """The batched form against the plain path run graph by graph on the same
    graphs: total arcs / wall seconds for each, and a bit-for-bit comparison of
    every batched result with its solo result."""
    import numpy as np
    import thegraphpartitioning_amd as tgp
    t0 = time.perf_counter()
    gs, ks = batched_workload(args.batched, SEED)
    arcs = sum(g.m for g in gs)
    print(f"# batched workload: {len(gs)} graphs, {sum(g.n for g in gs)} vertices, {arcs} arcs "
          f"(generated in {time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    dgs = [tgp.DeviceGraph(g) for g in gs]
    ptrs = [d.pointers() for d in dgs]
    kw = dict(init=tgp.INIT_BLOCK, repair=True)
    if args.batch_balance_eps > 0:
        kw["balance_eps"] = args.batch_balance_eps
    wn = min(len(gs), 2 * args.workers)
    tgp.partition_batch(gs[:wn], ks[:wn], workers=args.workers, device_graphs=ptrs[:wn], **kw)
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = tgp.partition_batch(gs, ks, workers=args.workers, device_graphs=ptrs, **kw)
        print(f"# batched step {i}: {time.perf_counter() - t0:.3f}s", file=sys.stderr, flush=True)
    b_s = (time.perf_counter() - t0) / args.steps
    # the plain path, graph by graph, on every graph or an evenly spaced sample
    step = max(1, len(gs) // args.plain_sample) if args.plain_sample else 1
    sample = list(range(0, len(gs), step))
    t0 = time.perf_counter()
    solo = {}
    for c, i in enumerate(sample):
        solo[i] = tgp.partition(gs[i], ks[i], device_graph=ptrs[i], **kw)
        if c % 64 == 63:
            print(f"# plain graph {c + 1}/{len(sample)}: {time.perf_counter() - t0:.3f}s",
                  file=sys.stderr, flush=True)
    p_s = time.perf_counter() - t0
    p_arcs = sum(gs[i].m for i in sample)
    same = sum(int(out[i][2] == solo[i][2] and np.array_equal(out[i][0], solo[i][0]) and
                   np.array_equal(out[i][1], solo[i][1])) for i in sample)
    for d in dgs:
        d.free()
    value = arcs / b_s / 1e6
    print(json.dumps({
        "metric": "M edges/s end-to-end (batched)", "value": round(value, 3), "unit": "M edges/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": round(b_s * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32+fp32",
        "data": "synthetic",
        "config": {"workload": f"batched: {len(gs)} graphs, n log-uniform in [1k, 200k], "
                               "half R-MAT ef 8, half Erdos-Renyi mean degree 10, k in {2,4,8,16}",
                   "graphs": len(gs), "arcs": arcs, "workers": args.workers,
                   "coarsening": "repaired", "initial_partition": "block i*k/n_c",
                   "balance_eps": args.batch_balance_eps or None,
                   "parallelism": "single GPU, graphs on concurrent streams"},
        "detail": {"plain_value": round(p_arcs / p_s / 1e6, 3),
                   "plain_graphs": len(sample), "plain_seconds": round(p_s, 3),
                   "speedup_vs_plain": round((p_arcs / p_s) and value / (p_arcs / p_s / 1e6), 3),
                   "bit_identical_to_solo": f"{same}/{len(sample)}"}}), flush=True)


def bench_batched_autotune(args):
    """The batched engine's tuner on a seeded random sample of the batched
    workload (both graph families): prints its measurements and writes the
    tuning table to --tune-out."""
    import numpy as np
    import thegraphpartitioning_amd as tgp
    gs, ks = batched_workload(args.batched, SEED)
    pick = np.random.default_rng(SEED + 7).choice(len(gs), min(args.batched_autotune, len(gs)),
                                                  replace=False)
    idx = sorted(int(i) for i in pick)
    sg, sk = [gs[i] for i in idx], [ks[i] for i in idx]
    print(f"# tuning sample: {len(sg)} of {len(gs)} graphs, {sum(g.m for g in sg)} arcs",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    report, _ = tgp.batch_autotune(sg, sk, path=args.tune_out)
    print(report, flush=True)
    print(f"# tuner: {time.perf_counter() - t0:.1f}s; table written to {args.tune_out}",
          file=sys.stderr, flush=True)


def bench_hbm_scale(args):
    """The HBM-scale configuration: R-MAT scale 26, edge factor 16, symmetrized
    (int64 offsets, ~2.1G arcs), int32 weights in [1, 16], k = 256, wide mode,
    repaired coarsening.  Too large for the oracle, so it reports throughput,
    per-family algorithmic GB/s, and the size-independent checks: a repeat run
    is identical, the reported cut equals a host recount, the part counts sum
    to n."""
    import numpy as np
    import thegraphpartitioning_amd as tgp
    t0 = time.perf_counter()
    g = tgp.gen_rmat(args.scale, args.edge_factor, SEED + 1, "hash16")
    print(f"# R-MAT {args.scale}/{args.edge_factor} hash16: n={g.n} m={g.m} "
          f"(generated in {time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    dg = tgp.DeviceGraph(g)
    kw = dict(init=tgp.INIT_BLOCK, repair=True, wide=True, device_graph=dg.pointers())
    a = tgp.partition(g, args.k, profile=True, **kw)
    st = a[3]
    fam = {name: {"ms": round(ms, 1), "GB/s": round(st.kernel_bytes[name] / ms / 1e6, 1)}
           for name, ms in st.kernel_ms.items() if ms > 1}
    print(f"# profiled partition: {st.seconds_total:.2f}s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = tgp.partition(g, args.k, **kw)
    secs = time.perf_counter() - t0
    part = b[0]
    recount = 0
    step = 1 << 22
    for v0 in range(0, g.n, step):
        v1 = min(g.n, v0 + step)
        e0, e1 = int(g.xadj[v0]), int(g.xadj[v1])
        src = np.repeat(np.arange(v0, v1, dtype=np.int64), np.diff(g.xadj[v0:v1 + 1]))
        diff = part[src] != part[g.adjncy[e0:e1]]
        recount += int(g.adjwgt[e0:e1][diff].astype(np.int64).sum())
    dg.free()
    checks = {"repeat_identical": bool(np.array_equal(a[0], b[0]) and a[2] == b[2] and
                                       np.array_equal(a[1], b[1])),
              "cut_equals_recount": b[2] == recount // 2,
              "counts_sum_to_n": int(b[1].sum()) == g.n,
              "parts_in_range": bool(part.min() >= 0 and part.max() < args.k)}
    print(json.dumps({
        "metric": "M edges/s end-to-end at k=256 (HBM-scale)", "value": round(g.m / secs / 1e6, 3),
        "unit": "M edges/s", "n_gpus": 1, "seconds": round(secs, 3), "dtype": "int32+fp32",
        "data": "synthetic",
        "config": {"workload": f"rmat{args.scale}-ef{args.edge_factor}-hash16-k{args.k}",
                   "vertices": g.n, "arcs": g.m, "k": args.k, "wide": True,
                   "coarsening": "repaired", "initial_partition": "block i*k/n_c"},
        "edge_cut": b[2], "levels": b[3].levels, "passes": b[3].passes, "checks": checks,
        "kernel_families": fam}), flush=True)
    if not all(checks.values()):
        sys.exit(1)


def bench_learned(args):
    """The learned move proposer against the plain path on the same GPU.

    Training: the plain partitioner's own refinement passes at the finest
    level on seeded R-MAT graphs (scales 14-16, edge factor 16, k as given;
    with more than 6 graphs, scales 16-20 streamed one graph at a time, a
    few evenly spaced passes each), or 3-D grids of side 24-48 with
    --learn-grid, one full-graph step per recorded pass.  Evaluation: R-MAT scale args.scale edge factor 16, or the
    --learn-grid grid in wide mode (resident in HBM), plain vs learned (learned
    passes at the finest level, every move admitted by its exact integer gain
    and the capacity rule, then the ordinary refinement): edges/s, edge cut
    and imbalance of each."""
    import numpy as np
    import torch
    import thegraphpartitioning_amd as tgp
    from thegraphpartitioning_amd import learn
    k = args.k
    grid = args.learn_grid
    t0 = time.perf_counter()
    model = learn.GNN(k, seed=0)
    if args.learn_graphs > 6 and not grid:
        # streamed: R-MAT scales 16-20 in turn, each graph's passes recorded,
        # args.learn_pairs of them (evenly spaced) trained on, then dropped
        from concurrent.futures import ThreadPoolExecutor
        npairs = 0
        loss = accu = 0.0
        steps = 0
        # graphs are generated on 8 host threads, 16 ahead of the GPU work
        pool = ThreadPoolExecutor(8)
        gen = lambda i: tgp.gen_rmat(16 + i % 5, 16, SEED + 100 + i)
        ahead = [pool.submit(gen, i) for i in range(min(16, args.learn_graphs))]
        for i in range(args.learn_graphs):
            g = ahead[i].result()
            ahead[i] = None
            if i + 16 < args.learn_graphs:
                ahead.append(pool.submit(gen, i + 16))
            ((_, pairs),) = learn.collect_traces([g], k, init=tgp.INIT_BLOCK, repair=True)
            if not pairs:
                continue
            sel = pairs[::max(1, len(pairs) // args.learn_pairs)][:args.learn_pairs]
            dg = learn.DevGraph(g)
            for _ in range(args.learn_epochs):
                for b, a in sel:
                    l, acc = model.train_step(dg, torch.from_numpy(b).cuda(),
                                              torch.from_numpy(a).cuda())
                    loss += l
                    accu += acc
                    steps += 1
            npairs += len(sel)
            del dg, pairs, sel
            if (i + 1) % 32 == 0:
                print(f"# {i + 1} graphs, {steps} steps: loss {loss / max(steps, 1):.4f} "
                      f"accuracy {accu / max(steps, 1):.4f} ({time.perf_counter() - t0:.1f}s)",
                      file=sys.stderr, flush=True)
                loss = accu = 0.0
                steps = 0
        pool.shutdown()
        ntrain = args.learn_graphs
    else:
        if grid:
            train = [tgp.gen_grid3d(s) for s in [24, 28, 32, 36, 40, 48][:args.learn_graphs]]
        else:
            train = [tgp.gen_rmat(s, 16, SEED + 100 + i) for i, s in
                     enumerate([14, 14, 15, 15, 16, 16][:args.learn_graphs])]
        data = learn.collect_traces(train, k, init=tgp.INIT_BLOCK, repair=True, wide=bool(grid))
        npairs = sum(len(p) for _, p in data)
        ntrain = len(train)
        print(f"# traces: {len(train)} graphs, {npairs} passes ({time.perf_counter() - t0:.1f}s)",
              file=sys.stderr, flush=True)
        dgs = [learn.DevGraph(g) for g, _ in data]
        items = [(i, j) for i, (_, p) in enumerate(data) for j in range(len(p))]
        rng = np.random.default_rng(0)
        for ep in range(args.learn_epochs):
            rng.shuffle(items)
            loss = accu = 0.0
            for i, j in items:
                b, a = data[i][1][j]
                l, acc = model.train_step(dgs[i], torch.from_numpy(b).cuda(),
                                          torch.from_numpy(a).cuda())
                loss += l
                accu += acc
            print(f"# epoch {ep}: loss {loss / len(items):.4f} accuracy {accu / len(items):.4f}",
                  file=sys.stderr, flush=True)
        del dgs, data
    train_s = time.perf_counter() - t0
    g = tgp.gen_grid3d(grid) if grid else tgp.gen_rmat(args.scale, args.edge_factor, SEED)
    dg = tgp.DeviceGraph(g)
    kw = dict(init=tgp.INIT_BLOCK, repair=True, wide=bool(grid), device_graph=dg.pointers())
    handle = model.handle(max_passes=20, min_admit=0.0005, margin=args.learn_margin,
                          infer_every=args.learn_infer_every)

    eps = args.balance_eps

    def run(learned, bal):
        extra = dict(learned=handle) if learned else {}
        if bal:
            extra["balance_eps"] = eps
        tgp.partition(g, k, **kw, **extra)  # warm-up
        t = time.perf_counter()
        r = tgp.partition(g, k, trace_cap=1 << 14, **kw, **extra)
        secs = time.perf_counter() - t
        part = r[0]
        imb = float(np.bincount(part, minlength=k).max()) / math.ceil(g.n / k) - 1.0
        out = dict(seconds=round(secs, 4), value=round(g.m / secs / 1e6, 3), edge_cut=r[2],
                   imbalance=round(imb, 4), passes=r[3].passes)
        rec = [t for t in r[3].trace if t[0] == 6]
        if rec:
            out["balance"] = {"rounds": rec[0][3], "balance_moves": rec[0][4],
                              "refine_moves": rec[0][5]}
        return out
    plain = run(False, False)
    plain_eps = run(False, True)
    lrn_eps = run(True, True)
    dg.free()
    print(json.dumps({
        "metric": f"M edges/s end-to-end at k={k} (learned, eps={eps})", "value": lrn_eps["value"],
        "unit": "M edges/s", "n_gpus": 1, "dtype": "bf16+fp32", "data": "synthetic",
        "config": {"workload": f"grid{grid}^3-k{k}" if grid else
                   f"rmat{args.scale}-ef{args.edge_factor}-k{k}", "vertices": g.n,
                   "arcs": g.m, "train_graphs": ntrain, "train_passes": npairs,
                   "train_epochs": args.learn_epochs, "train_seconds": round(train_s, 2),
                   "coarsening": "repaired", "initial_partition": "block i*k/n_c",
                   "balance_eps": eps, "infer_every": args.learn_infer_every},
        "plain": plain, "plain_eps": plain_eps, "learned_eps": lrn_eps,
        "cut_vs_plain_eps": round(lrn_eps["edge_cut"] / max(plain_eps["edge_cut"], 1), 4),
        "speed_vs_plain_eps": round(lrn_eps["value"] / plain_eps["value"], 4),
        "speed_vs_plain": round(lrn_eps["value"] / plain["value"], 4)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-scale", type=int, default=20,
                    help="R-MAT scale of the oracle's all-threads CPU baseline run")
    ap.add_argument("--cpu-scale-one", type=int, default=18,
                    help="R-MAT scale of the oracle's one-thread CPU baseline run")
    ap.add_argument("--profile-only", action="store_true",
                    help="run one profiled step and print per-kernel times")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (counter calibration)")
    ap.add_argument("--batched", type=int, default=0, metavar="GRAPHS",
                    help="measure the batched form on GRAPHS independent graphs instead")
    ap.add_argument("--workers", type=int, default=16,
                    help="graphs partitioned at once by the batched form's stream workers")
    ap.add_argument("--learned", action="store_true",
                    help="train the learned move proposer and compare it with the plain path")
    ap.add_argument("--learn-graphs", type=int, default=6,
                    help="with --learned: training graphs; above 6, R-MAT scales 16-20 "
                         "streamed one graph at a time")
    ap.add_argument("--learn-pairs", type=int, default=4,
                    help="recorded passes trained on per streamed graph")
    ap.add_argument("--learn-epochs", type=int, default=3)
    ap.add_argument("--learn-grid", type=int, default=0, metavar="SIDE",
                    help="with --learned: train on 3-D grids and evaluate on the SIDE^3 grid "
                         "(wide mode) instead of R-MAT")
    ap.add_argument("--balance-eps", type=float, default=0.03,
                    help="imbalance bound of the balance stage in the learned comparison")
    ap.add_argument("--batched-autotune", type=int, default=0, metavar="GRAPHS",
                    help="with --batched: run the engine's tuner on a random sample of this "
                         "many graphs of the workload and write its table")
    ap.add_argument("--tune-out", default="batch_tune.txt",
                    help="where --batched-autotune writes the tuning table")
    ap.add_argument("--batch-balance-eps", type=float, default=0.0,
                    help="with --batched: run every graph with this balance bound "
                         "(tgp_options.balance_eps; 0 = the reference's pipeline)")
    ap.add_argument("--learn-infer-every", type=int, default=4,
                    help="the GNN runs every this many learned passes (logits reused between)")
    ap.add_argument("--learn-margin", type=float, default=3.0,
                    help="a proposal's logit may trail the own part's by this much")
    ap.add_argument("--hbm-scale", action="store_true",
                    help="run the HBM-scale configuration (use with --scale 26 --k 256)")
    ap.add_argument("--plain-sample", type=int, default=0, metavar="GRAPHS",
                    help="time the plain path on an evenly spaced sample of this many graphs "
                         "(0: all of them)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described, not built)", file=sys.stderr)
        sys.exit(2)
    if args.batched and args.batched_autotune:
        bench_batched_autotune(args)
        return
    if args.batched:
        bench_batched(args)
        return
    if args.hbm_scale:
        bench_hbm_scale(args)
        return
    if args.learned:
        bench_learned(args)
        return

    import thegraphpartitioning_amd as tgp
    t0 = time.perf_counter()
    g = tgp.gen_rmat(args.scale, args.edge_factor, SEED)
    gen_s = time.perf_counter() - t0
    print(f"# graph R-MAT {args.scale}/{args.edge_factor}: n={g.n} m={g.m} "
          f"(generated in {gen_s:.1f}s)", file=sys.stderr, flush=True)
    dg = tgp.DeviceGraph(g)
    kw = dict(init=tgp.INIT_BLOCK, repair=True, device_graph=dg.pointers())
    if args.profile_only:
        rp = tgp.partition(g, args.k, profile=True, **kw)
        st = rp[3]
        print(json.dumps({"n": g.n, "m": g.m, "seconds": st.seconds_total, "levels": st.levels,
                          "passes": st.passes, "match_rounds": st.match_rounds,
                          "sorts_skipped": st.sorts_skipped,
                          "cut": rp[2], "coarsest_n": st.coarsest_n,
                          "kernel_ms": {a: round(b, 2) for a, b in st.kernel_ms.items()},
                          "launches": st.kernel_launches}), flush=True)
        dg.free()
        return

    for i in range(args.warmup):
        r = tgp.partition(g, args.k, **kw)
        print(f"# warmup {i}: cut={r[2]} levels={r[3].levels} passes={r[3].passes} "
              f"{r[3].seconds_total:.3f}s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = tgp.partition(g, args.k, **kw)  # synchronous: returns after D2H + stream sync
        print(f"# step {i}: {time.perf_counter() - t0:.3f}s", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    part, pc, cut, st = r
    value = g.m * args.steps / elapsed / 1e6
    print(f"# timed {args.steps} steps: {elapsed:.3f}s cut={cut}", file=sys.stderr, flush=True)

    # one untimed step from host buffers: the rate including the CSR upload over PCIe
    kw_host = dict(init=tgp.INIT_BLOCK, repair=True)
    t2 = time.perf_counter()
    tgp.partition(g, args.k, **kw_host)
    pcie_value = g.m / (time.perf_counter() - t2) / 1e6

    # one extra step with per-kernel HIP-event timing (not part of the timed region)
    rp = tgp.partition(g, args.k, profile=True, trace_cap=1 << 14, **kw)
    stp = rp[3]
    total_kernel_ms = sum(stp.kernel_ms.values())

    def kernel_line(fam):
        ms = stp.kernel_ms[fam]
        byts = stp.kernel_bytes[fam]
        launches = stp.kernel_launches[fam]
        achieved = byts / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        kernel = ROOFLINE_FAMILIES[fam]
        return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(kernel), "launches": launches,
                "avg_launch_ms": round(ms / max(launches, 1), 4),
                "bytes_per_launch": round(byts / max(launches, 1)),
                "kernel_share_of_step": round(ms / max(total_kernel_ms, 1e-9), 3)}
    fams = sorted(ROOFLINE_FAMILIES, key=lambda f: -stp.kernel_ms.get(f, 0.0))
    import numpy as np
    sizes = np.bincount(part[part >= 0], minlength=args.k)
    detail = {
        "levels": st.levels, "passes": st.passes, "coarsest_n": st.coarsest_n,
        "edge_cut": int(cut), "n": g.n, "m": g.m,
        "imbalance": round(float(sizes.max()) / ((g.n + args.k - 1) // args.k) - 1.0, 4),
        "kernel_ms": {k_: round(v, 3) for k_, v in stp.kernel_ms.items() if v > 0},
        "kernel_launches": {k_: v for k_, v in stp.kernel_launches.items() if v > 0},
        "kernel_gbs": {k_: round(stp.kernel_bytes[k_] / (v / 1e3) / 1e9, 1)
                       for k_, v in stp.kernel_ms.items() if v > 0},
        "kernel_ms_total": round(total_kernel_ms, 3),
        "profiled_step_s": round(stp.seconds_total, 4),
        "match_launches": stp.match_rounds,
        "sorts_skipped": stp.sorts_skipped,
        "pcie_inclusive_value": round(pcie_value, 3),
    }
    # the whole step against the same peak: the per-kernel model's bytes of
    # every level and pass this partition ran, over the timed step
    e2e_bytes = model_bytes(stp.trace)
    e2e_gbs = e2e_bytes / (elapsed / args.steps) / 1e9
    roofline = kernel_line(fams[0])
    roofline.update({"e2e_model_bytes": round(e2e_bytes), "e2e_achieved": round(e2e_gbs, 1),
                     "e2e_frac": round(e2e_gbs / HBM_PEAK_GBS, 4)})
    roofline["next_kernel"] = kernel_line(fams[1])
    dg.free()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_scale, args.edge_factor, args.k, args.cpu_scale_one)

    out = {
        "metric": "M edges/s end-to-end at k=64 (plain)",
        "value": round(value, 3),
        "unit": "M edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic",
        "config": {"workload": f"rmat{args.scale}-ef{args.edge_factor}-k{args.k}",
                   "vertices": g.n, "arcs": g.m, "k": args.k,
                   "coarsening": "repaired (reference coarsening stalls on this graph)",
                   "initial_partition": "block i*k/n_c (no METIS call)",
                   "input": "CSR resident in HBM when the timed region starts",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": detail,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
