"""Learned move proposer: a 2-layer mean-aggregation GNN trained on the plain
refinement's own passes (no counterpart in the reference).

The arithmetic runs in the native kernels of include/tgp_learn.h (BF16 MFMA
GEMMs with FP32 accumulation, mean aggregation, features, loss, Adam); torch
only holds the device buffers.

    model = GNN(k)
    data = collect_traces(graphs, k)            # plain refinement, finest level
    model.fit(data, epochs=...)
    part, pc, cut, st = tgp.partition(g, k, learned=model.handle(), ...)

Inputs per vertex: x_v = [T_v[q]/deg_v (q < k), onehot(part_v), log2(1+deg_v)/16,
1] (padded to a multiple of 32 columns); hidden 64; output k logits.  The first
layer's mean term is computed as mean_N(X W1b) (64 columns) rather than
mean_N(X) W1b (F columns); the two are equal in exact arithmetic.  The
label of a vertex in a pass is the part it holds after the plain pass.
"""
import ctypes as C
import math

import numpy as np

import thegraphpartitioning_amd as tgp

HIDDEN = 64


class CModel(C.Structure):
    _fields_ = [("k", C.c_int32), ("features", C.c_int32), ("hidden", C.c_int32),
                ("max_passes", C.c_int32), ("W1a", C.c_void_p), ("W1b", C.c_void_p),
                ("W2a", C.c_void_p), ("W2b", C.c_void_p), ("b1", C.c_void_p),
                ("b2", C.c_void_p), ("min_admit", C.c_float), ("margin", C.c_float),
                ("infer_every", C.c_int32)]


def _lib():
    L = tgp.lib()
    if getattr(L, "_learn_ready", False):
        return L
    P, I, I64, F32 = C.c_void_p, C.c_int, C.c_int64, C.c_float
    L.tgpl_gemm.argtypes = [P, I, I, P, I, I, P, I, P, I, I, I, F32, F32, P, I, P]
    L.tgpl_gemm_ws.argtypes = L.tgpl_gemm.argtypes[:-1] + [P, I64, P]
    L.tgpl_mean_agg.argtypes = [I64, P, P, P, I, I, P, P, P]
    L.tgpl_features.argtypes = [I64, P, P, P, P, I, I, P, P, P]
    L.tgpl_softmax_xent.argtypes = [I64, I, P, P, P, F32, F32, F32, P, P, P, P]
    L.tgpl_relu_bwd.argtypes = [I64, P, P, P, P]
    L.tgpl_colsum.argtypes = [I64, I, P, P, P]
    L.tgpl_adam.argtypes = [I64, P, P, P, P, F32, F32, F32, F32, I, P, P]
    L.tgpl_to_bf16.argtypes = [I64, P, P, P]
    L.tgpl_axpy.argtypes = [I64, F32, P, P, P]
    L.tgpl_propose.argtypes = [I64, I, P, P, P, F32, P, P, P]
    for name in ("tgpl_gemm", "tgpl_mean_agg", "tgpl_features", "tgpl_softmax_xent",
                 "tgpl_relu_bwd", "tgpl_colsum", "tgpl_adam", "tgpl_to_bf16", "tgpl_axpy",
                 "tgpl_propose"):
        getattr(L, name).restype = C.c_int
    L._learn_ready = True
    return L


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def feature_count(k):
    return ((2 * k + 2 + 31) // 32) * 32


class DevGraph:
    """A CSR in device memory (torch tensors) for the learning kernels."""

    def __init__(self, g, device="cuda"):
        import torch
        self.n, self.m = g.n, g.m
        self.xadj = torch.from_numpy(np.ascontiguousarray(g.xadj)).to(device)
        self.adj = torch.from_numpy(np.ascontiguousarray(g.adjncy)).to(device)
        self.w = torch.from_numpy(np.ascontiguousarray(g.adjwgt)).to(device)


class GNN:
    """2-layer mean-aggregation GNN, FP32 master weights with BF16 copies."""

    def __init__(self, k, hidden=HIDDEN, seed=0, device="cuda"):
        import torch
        self.torch = torch
        self.k, self.hidden = k, hidden
        self.F = feature_count(k)
        self.device = device
        gen = torch.Generator().manual_seed(seed)

        def glorot(r, c):
            a = math.sqrt(6.0 / (r + c))
            return ((torch.rand(r, c, generator=gen) * 2 - 1) * a).to(device)
        self.W = {"W1a": glorot(self.F, hidden), "W1b": glorot(self.F, hidden),
                  "b1": torch.zeros(hidden, device=device),
                  "W2a": glorot(hidden, k), "W2b": glorot(hidden, k),
                  "b2": torch.zeros(k, device=device)}
        self.Wb = {n: torch.empty(w.shape, dtype=torch.bfloat16, device=device)
                   for n, w in self.W.items() if n[0] == "W"}
        self.m = {n: torch.zeros_like(w) for n, w in self.W.items()}
        # split-K workspace of the weight-gradient products (K = vertices)
        self.work = torch.empty(8 << 20, device=device)
        self.v = {n: torch.zeros_like(w) for n, w in self.W.items()}
        self.step = 0
        self._refresh_bf16()

    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def _refresh_bf16(self):
        L = _lib()
        for n, wb in self.Wb.items():
            _ok(L.tgpl_to_bf16(self.W[n].numel(), _p(self.W[n]), _p(wb), self._stream()),
                "to_bf16")

    # ---- forward ------------------------------------------------------------
    def forward(self, dg, part):
        """Logits (n x k, FP32) for partition `part` (int32 device tensor);
        also the intermediates the backward pass needs."""
        torch, L, s = self.torch, _lib(), self._stream()
        n, F, Hd, k = dg.n, self.F, self.hidden, self.k
        bf = torch.bfloat16
        X = torch.empty(n, F, dtype=bf, device=self.device)
        tally = torch.empty(n, k, dtype=torch.int32, device=self.device)
        _ok(L.tgpl_features(n, _p(dg.xadj), _p(dg.adj), _p(dg.w), _p(part), k, F, _p(X),
                            _p(tally), s), "features")
        # mean_N(X) W1b = mean_N(X W1b): aggregate the 64 columns of P = X W1b
        P = torch.empty(n, Hd, device=self.device)
        Pb = torch.empty(n, Hd, dtype=bf, device=self.device)
        _ok(L.tgpl_gemm(_p(X), F, 0, _p(self.Wb["W1b"]), Hd, 0, _p(P), Hd, _p(Pb), n, Hd, F,
                        1.0, 0.0, None, 0, s), "gemm")
        H = torch.empty(n, Hd, device=self.device)  # starts as mean_N(P)
        M0b = torch.empty(n, Hd, dtype=bf, device=self.device)
        _ok(L.tgpl_mean_agg(n, _p(dg.xadj), _p(dg.adj), _p(Pb), Hd, 0, _p(M0b), _p(H), s), "agg")
        Hb = torch.empty(n, Hd, dtype=bf, device=self.device)
        _ok(L.tgpl_gemm(_p(X), F, 0, _p(self.Wb["W1a"]), Hd, 0, _p(H), Hd, _p(Hb), n, Hd, F,
                        1.0, 1.0, _p(self.W["b1"]), 1, s), "gemm")
        M1 = torch.empty(n, Hd, dtype=bf, device=self.device)
        _ok(L.tgpl_mean_agg(n, _p(dg.xadj), _p(dg.adj), _p(Hb), Hd, 0, _p(M1), None, s), "agg")
        Z = torch.empty(n, k, device=self.device)
        _ok(L.tgpl_gemm(_p(Hb), Hd, 0, _p(self.Wb["W2a"]), k, 0, _p(Z), k, None, n, k, Hd,
                        1.0, 0.0, None, 0, s), "gemm")
        _ok(L.tgpl_gemm(_p(M1), Hd, 0, _p(self.Wb["W2b"]), k, 0, _p(Z), k, None, n, k, Hd,
                        1.0, 1.0, _p(self.W["b2"]), 0, s), "gemm")
        return Z, dict(X=X, H=H, Hb=Hb, M1=M1, tally=tally)

    # ---- one training step ---------------------------------------------------
    def train_step(self, dg, part, label, lr=3e-3, wmove=8.0):
        """Forward, weighted cross-entropy against `label`, backward, Adam.
        Returns (mean loss, accuracy)."""
        torch, L, s = self.torch, _lib(), self._stream()
        n, F, Hd, k = dg.n, self.F, self.hidden, self.k
        bf = torch.bfloat16
        Z, c = self.forward(dg, part)
        nmove = int((label != part).sum().item())
        wsum = (n - nmove) * 1.0 + nmove * wmove
        dZ = torch.empty(n, k, device=self.device)
        dZb = torch.empty(n, k, dtype=bf, device=self.device)
        acc = torch.zeros(4, dtype=torch.float64, device=self.device)
        _ok(L.tgpl_softmax_xent(n, k, _p(Z), _p(label), _p(part), 1.0, wmove, 1.0 / wsum,
                                _p(dZ), _p(dZb), _p(acc), s), "xent")
        G = {}
        G["W2a"] = torch.empty(Hd, k, device=self.device)
        G["W2b"] = torch.empty(Hd, k, device=self.device)
        ws, wn = _p(self.work), self.work.numel()
        _ok(L.tgpl_gemm_ws(_p(c["Hb"]), Hd, 1, _p(dZb), k, 0, _p(G["W2a"]), k, None, Hd, k, n,
                           1.0, 0.0, None, 0, ws, wn, s), "gemm")
        _ok(L.tgpl_gemm_ws(_p(c["M1"]), Hd, 1, _p(dZb), k, 0, _p(G["W2b"]), k, None, Hd, k, n,
                           1.0, 0.0, None, 0, ws, wn, s), "gemm")
        G["b2"] = torch.empty(k, device=self.device)
        _ok(L.tgpl_colsum(n, k, _p(dZ), _p(G["b2"]), s), "colsum")
        dH = torch.empty(n, Hd, device=self.device)
        _ok(L.tgpl_gemm(_p(dZb), k, 0, _p(self.Wb["W2a"]), k, 1, _p(dH), Hd, None, n, Hd, k,
                        1.0, 0.0, None, 0, s), "gemm")
        dM1 = torch.empty(n, Hd, device=self.device)
        dM1b = torch.empty(n, Hd, dtype=bf, device=self.device)
        _ok(L.tgpl_gemm(_p(dZb), k, 0, _p(self.Wb["W2b"]), k, 1, _p(dM1), Hd, _p(dM1b), n, Hd,
                        k, 1.0, 0.0, None, 0, s), "gemm")
        back = torch.empty(n, Hd, device=self.device)
        backb = torch.empty(n, Hd, dtype=bf, device=self.device)
        _ok(L.tgpl_mean_agg(n, _p(dg.xadj), _p(dg.adj), _p(dM1b), Hd, 1, _p(backb), _p(back), s),
            "agg")
        _ok(L.tgpl_axpy(n * Hd, 1.0, _p(back), _p(dH), s), "axpy")
        dHb = torch.empty(n, Hd, dtype=bf, device=self.device)
        _ok(L.tgpl_relu_bwd(n * Hd, _p(c["H"]), _p(dH), _p(dHb), s), "relu")
        G["W1a"] = torch.empty(F, Hd, device=self.device)
        G["W1b"] = torch.empty(F, Hd, device=self.device)
        _ok(L.tgpl_gemm_ws(_p(c["X"]), F, 1, _p(dHb), Hd, 0, _p(G["W1a"]), Hd, None, F, Hd, n,
                           1.0, 0.0, None, 0, ws, wn, s), "gemm")
        # H's mean term: dP = adjoint mean of dH, dW1b = X^T dP
        dPf = torch.empty(n, Hd, device=self.device)
        dPb = torch.empty(n, Hd, dtype=bf, device=self.device)
        _ok(L.tgpl_mean_agg(n, _p(dg.xadj), _p(dg.adj), _p(dHb), Hd, 1, _p(dPb), _p(dPf), s),
            "agg")
        _ok(L.tgpl_gemm_ws(_p(c["X"]), F, 1, _p(dPb), Hd, 0, _p(G["W1b"]), Hd, None, F, Hd, n,
                           1.0, 0.0, None, 0, ws, wn, s), "gemm")
        G["b1"] = torch.empty(Hd, device=self.device)
        _ok(L.tgpl_colsum(n, Hd, _p(dH), _p(G["b1"]), s), "colsum")
        self.step += 1
        for name, w in self.W.items():
            _ok(L.tgpl_adam(w.numel(), _p(w), _p(G[name]), _p(self.m[name]), _p(self.v[name]),
                            lr, 0.9, 0.999, 1e-8, self.step, _p(self.Wb.get(name)), s), "adam")
        a = acc.cpu().numpy()
        return a[0] / max(a[1], 1e-30), a[2] / max(a[3], 1)

    def evaluate(self, dg, part, label):
        """(mean weighted loss, accuracy, recall of moved vertices)."""
        torch, L, s = self.torch, _lib(), self._stream()
        Z, _ = self.forward(dg, part)
        acc = torch.zeros(4, dtype=torch.float64, device=self.device)
        _ok(L.tgpl_softmax_xent(dg.n, self.k, _p(Z), _p(label), _p(part), 1.0, 1.0, 1.0, None,
                                None, _p(acc), s), "xent")
        moved = label != part
        pred = Z.argmax(1).to(torch.int32)
        rec = float(((pred == label) & moved).sum().item()) / max(int(moved.sum().item()), 1)
        a = acc.cpu().numpy()
        return a[0] / max(a[1], 1e-30), a[2] / max(a[3], 1), rec

    # ---- use in the partitioner ------------------------------------------------
    def handle(self, max_passes=20, min_admit=0.0005, margin=3.0, infer_every=1):
        """A tgp_learned_model for tgp.partition(..., learned=...); keep self alive.

        The library reads the weights on its own stream, so the work that
        wrote them on torch's current stream (training, load_state) is
        finished here first."""
        self.torch.cuda.current_stream().synchronize()
        self._cm = CModel(self.k, self.F, self.hidden, max_passes,
                          self.Wb["W1a"].data_ptr(), self.Wb["W1b"].data_ptr(),
                          self.Wb["W2a"].data_ptr(), self.Wb["W2b"].data_ptr(),
                          self.W["b1"].data_ptr(), self.W["b2"].data_ptr(), min_admit, margin,
                          infer_every)
        return self._cm

    def state(self):
        return {n: w.cpu().numpy() for n, w in self.W.items()}

    def load_state(self, st):
        torch = self.torch
        for n in self.W:
            self.W[n] = torch.from_numpy(np.asarray(st[n], np.float32)).to(self.device)
        self._refresh_bf16()

    def save(self, path):
        np.savez(path, k=self.k, hidden=self.hidden, **self.state())

    @classmethod
    def load(cls, path, device="cuda"):
        d = np.load(path)
        m = cls(int(d["k"]), int(d["hidden"]), device=device)
        m.load_state({n: d[n] for n in m.W})
        return m


def collect_traces(graphs, k, passes_per_graph=None, **kw):
    """Runs the plain partitioner on each graph and records, at the finest
    level, (partition before, partition after) for every refinement pass."""
    out = []
    for g in graphs:
        seq = []

        def dump(level, phase, pss, n, part, user):
            if level == 0:
                seq.append((phase, pss, np.ctypeslib.as_array(part, (n,)).copy()))
        fn = tgp.PASS_DUMP_FN(dump)
        tgp.partition(g, k, pass_dump=fn, **kw)
        pairs = []
        for a, b in zip(seq, seq[1:]):
            if b[1] >= 0 and a[0] == b[0]:  # consecutive states of one phase
                if not np.array_equal(a[2], b[2]):
                    pairs.append((a[2], b[2]))
        if passes_per_graph:
            pairs = pairs[:passes_per_graph]
        out.append((g, pairs))
    return out
