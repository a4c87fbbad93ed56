"""Seeded test graphs shared by the CPU and GPU test files."""
import numpy as np


def csr_from_edges(n, edges, sym=True, weights=None, sort_rows=True):
    """CSR from an edge list; duplicates kept, self loops kept as given."""
    rows = [[] for _ in range(n)]
    for idx, (u, v) in enumerate(edges):
        w = 1 if weights is None else int(weights[idx])
        rows[u].append((v, w))
        if sym and u != v:
            rows[v].append((u, w))
    xadj = [0]
    adj, wt = [], []
    for r in rows:
        if sort_rows:
            r = sorted(r)
        for v, w in r:
            adj.append(v)
            wt.append(w)
        xadj.append(len(adj))
    return (np.array(xadj, np.int64), np.array(adj, np.int32), np.array(wt, np.int32))


def readme_graph():
    """The 5-vertex CSR of the reference README (README.md:58-82)."""
    return (np.array([0, 2, 4, 4, 5, 6], np.int64), np.array([1, 2, 0, 3, 4, 2], np.int32),
            np.ones(6, np.int32))


def path_graph(n):
    return csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])


def random_graph(n, m, seed, sym=True, wmax=1, dup=False, selfloops=False):
    rng = np.random.default_rng(seed)
    edges = set() if not dup else []
    tries = 0
    while len(edges) < m and tries < 20 * m:
        tries += 1
        u, v = int(rng.integers(n)), int(rng.integers(n))
        if u == v and not selfloops:
            continue
        if dup:
            edges.append((u, v))
        else:
            edges.add((min(u, v), max(u, v)) if sym else (u, v))
    edges = list(edges) if not dup else edges
    w = rng.integers(1, wmax + 1, size=len(edges)) if wmax > 1 else None
    xadj, adj, wt = csr_from_edges(n, edges, sym=sym, weights=w, sort_rows=not dup)
    return xadj, adj, wt


def hub_graph(n, hubs, extra, seed, wmax=1, wbig=0):
    """Vertices 0..hubs-1 are hubs joined to every vertex (rows of n-1 arcs);
    `extra` random edges among the others; weights 1..wmax, or wbig on hub
    arcs when given (so a hub row's sum of |w| can pass 2**24)."""
    rng = np.random.default_rng(seed)
    edges = set()
    for h in range(hubs):
        for v in range(h + 1, n):
            edges.add((h, v))
    while len(edges) < extra + sum(n - 1 - h for h in range(hubs)):
        u, v = int(rng.integers(hubs, n)), int(rng.integers(hubs, n))
        if u != v:
            edges.add((min(u, v), max(u, v)))
    edges = sorted(edges)
    w = rng.integers(1, wmax + 1, size=len(edges))
    if wbig:
        w = np.where(np.array([e[0] < hubs for e in edges]), wbig, w)
    return csr_from_edges(n, edges, weights=w)


def cut_recount(g, part, step=1 << 22):
    """Edge cut of `part` over all arcs of g (sum of w over arcs whose ends lie
    in different parts, halved), with torch on the GPU in vertex chunks (an
    independent count: no code of the partitioner runs)."""
    import torch
    import thegraphpartitioning_amd as tgp
    tgp.release_device_memory()          # the library's cached blocks go back first
    dev = "cuda"
    p = torch.from_numpy(np.ascontiguousarray(part)).to(dev)
    total = 0
    for v0 in range(0, g.n, step):
        v1 = min(g.n, v0 + step)
        e0, e1 = int(g.xadj[v0]), int(g.xadj[v1])
        deg = torch.from_numpy(np.diff(g.xadj[v0:v1 + 1])).to(dev)
        src = torch.repeat_interleave(torch.arange(v0, v1, device=dev), deg)
        adj = torch.from_numpy(g.adjncy[e0:e1]).to(dev).long()
        w = torch.from_numpy(g.adjwgt[e0:e1]).to(dev).long()
        total += int(torch.where(p[src] != p[adj], w, torch.zeros_like(w)).sum().item())
        del deg, src, adj, w
    del p
    torch.cuda.empty_cache()
    return total // 2
