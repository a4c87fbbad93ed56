"""MI355X multilevel k-way graph partitioner (ctypes binding of libtgp.so).

A drop-in for the partitioning path of iamrupamroyy/TheGraphPartitioning
(src/gpu_part.cu): heavy-edge matching, contraction, an initial partition of
the coarsest graph, then projection and deterministic refinement, all on one
GPU through hand-written gfx950 kernels.  The `gpu_part` command built next to
this file keeps the reference's command line and files; this module exposes
the same path to Python callers.

Every call goes to the native library.  There is no CPU fallback: a missing
library raises ImportError and a GPU failure raises TGPError.
"""
import ctypes as C
import os

import numpy as np

__all__ = ["partition", "gen_rmat", "gen_grid3d", "gen_er", "Graph", "TGPError", "Stats",
           "read_graph_dir", "write_final_partition", "metis_text", "lib", "library_path",
           "KERNEL_NAMES", "INIT_GPMETIS", "INIT_BLOCK"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtgp.so")
CLI_PATH = os.path.join(_HERE, "gpu_part")

# size limits of the entry point that allocates nothing (tgp.h)
NOALLOC_MAX_N = 1 << 20
NOALLOC_MAX_M = 1 << 23
NOALLOC_MAX_K = 256

INIT_GPMETIS = "gpmetis"
INIT_BLOCK = "block"

TGP_OK, TGP_EINVAL, TGP_ELIMIT, TGP_EEDGES = 0, -1, -2, -3
TGP_EINITPART, TGP_EPARTFILE, TGP_EHIP, TGP_ENOMEM, TGP_EIO = -4, -5, -6, -7, -8
NKERNELS = 16


class CGraph(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("xadj", C.c_void_p),
                ("adjncy", C.c_void_p), ("adjwgt", C.c_void_p)]


INITPART_FN = C.CFUNCTYPE(C.c_int, C.POINTER(CGraph), C.c_int64, C.c_int32,
                          C.POINTER(C.c_int32), C.c_void_p)
PASS_DUMP_FN = C.CFUNCTYPE(None, C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                           C.POINTER(C.c_int32), C.c_void_p)


class COptions(C.Structure):
    _fields_ = [("k", C.c_int32), ("wide", C.c_int32), ("coarsen_floor", C.c_int64),
                ("initpart", C.c_void_p), ("initpart_user", C.c_void_p),
                ("device_input", C.c_int32), ("profile", C.c_int32),
                ("check_level", C.c_int32), ("repair", C.c_int32),
                ("pass_dump", C.c_void_p), ("pass_dump_user", C.c_void_p),
                ("learned", C.c_void_p), ("balance_eps", C.c_float)]


class CGpmetisDir(C.Structure):
    """tgp_gpmetis_dir: the working directory of one gpmetis call."""
    _fields_ = [("dir", C.c_char_p), ("status", C.c_int32), ("reserved", C.c_int32)]


class CBatchItem(C.Structure):
    pass


class CTrace(C.Structure):
    _fields_ = [("kind", C.c_int32), ("level", C.c_int32), ("phase", C.c_int32),
                ("pass_", C.c_int32), ("a", C.c_int64), ("b", C.c_int64), ("c", C.c_int64),
                ("d", C.c_int64), ("hash", C.c_uint64)]


class CStats(C.Structure):
    _fields_ = [("levels", C.c_int32), ("passes", C.c_int32), ("coarsest_n", C.c_int64),
                ("coarsest_npairs", C.c_int64), ("match_rounds", C.c_int64),
                ("seconds_total", C.c_double), ("seconds_initpart", C.c_double),
                ("kernel_ms", C.c_double * NKERNELS),
                ("kernel_launches", C.c_int64 * NKERNELS),
                ("kernel_bytes", C.c_double * NKERNELS),
                ("trace", C.POINTER(CTrace)), ("trace_cap", C.c_int64),
                ("trace_len", C.c_int64), ("fail_level", C.c_int64),
                ("fail_value", C.c_int64), ("sorts_skipped", C.c_int64),
                ("level0_form", C.c_int64)]


_lib = None


CBatchItem._fields_ = [("g", CGraph), ("o", COptions), ("part", C.c_void_p),
                       ("part_count", C.c_void_p), ("stats", C.POINTER(CStats)),
                       ("edge_cut", C.c_int64), ("rc", C.c_int32), ("reserved", C.c_int32)]


def library_path():
    return _LIB_PATH


def lib():
    """Loads libtgp.so (built in-tree by thegraphpartitioning_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError("libtgp.so is not built; run `python -m thegraphpartitioning_amd.build`")
    L = C.CDLL(_LIB_PATH)
    P = C.c_void_p
    L.tgp_partition.argtypes = [C.POINTER(CGraph), C.POINTER(COptions), P, P, P,
                                C.POINTER(CStats)]
    L.tgp_partition.restype = C.c_int
    L.tgp_partition_noalloc.argtypes = L.tgp_partition.argtypes
    L.tgp_partition_noalloc.restype = C.c_int
    L.tgp_balance.argtypes = L.tgp_partition.argtypes
    L.tgp_balance.restype = C.c_int
    L.tgp_batch_tuning.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    L.tgp_batch_tuning.restype = C.c_int
    L.tgp_batch_autotune.argtypes = [C.c_void_p, C.c_int64, C.c_char_p, C.c_void_p]
    L.tgp_batch_autotune.restype = C.c_int
    L.tgp_strerror.argtypes = [C.c_int]
    L.tgp_strerror.restype = C.c_char_p
    L.tgp_last_error.restype = C.c_char_p
    L.tgp_kernel_name.argtypes = [C.c_int]
    L.tgp_kernel_name.restype = C.c_char_p
    for name in ("tgp_gen_rmat", "tgp_gen_grid3d", "tgp_gen_er", "tgp_read_graph_dir"):
        getattr(L, name).restype = C.c_int
    L.tgp_gen_rmat.argtypes = [C.c_int32, C.c_int32, C.c_uint64, C.c_int32, C.POINTER(CGraph)]
    L.tgp_gen_grid3d.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(CGraph)]
    L.tgp_gen_er.argtypes = [C.c_int64, C.c_double, C.c_uint64, C.POINTER(CGraph)]
    L.tgp_gen_rmat_n.argtypes = [C.c_int64, C.c_int32, C.c_uint64, C.POINTER(CGraph)]
    L.tgp_read_csr_files.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(CGraph)]
    L.tgp_read_csr_files.restype = C.c_int
    L.tgp_write_bcsr.argtypes = [C.c_char_p, C.POINTER(CGraph)]
    L.tgp_write_bcsr.restype = C.c_int
    L.tgp_read_bcsr.argtypes = [C.c_char_p, C.POINTER(CGraph)]
    L.tgp_read_bcsr.restype = C.c_int
    L.tgp_write_part_file.argtypes = [C.c_char_p, P, C.c_int64]
    L.tgp_write_part_file.restype = C.c_int
    L.tgp_convert_edges.argtypes = [C.c_char_p, C.POINTER(CGraph)]
    L.tgp_convert_edges.restype = C.c_int
    L.tgp_write_conversion_files.argtypes = [C.POINTER(CGraph), C.c_char_p]
    L.tgp_write_conversion_files.restype = C.c_int
    L.tgp_partition_batch.argtypes = [C.POINTER(CBatchItem), C.c_int64, C.c_int32]
    L.tgp_partition_batch.restype = C.c_int
    L.tgp_read_graph_dir.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.POINTER(CGraph)]
    L.tgp_graph_free.argtypes = [C.POINTER(CGraph)]
    L.tgp_graph_free.restype = None
    L.tgp_write_final_partition.argtypes = [C.c_char_p, P, C.c_int64, C.c_int32]
    L.tgp_write_final_partition.restype = C.c_int
    L.tgp_metis_text.argtypes = [C.POINTER(CGraph), C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.tgp_metis_text.restype = C.c_int
    L.tgp_free.argtypes = [P]
    L.tgp_free.restype = None
    L.tgp_initpart_block.argtypes = [C.POINTER(CGraph), C.c_int64, C.c_int32,
                                     C.POINTER(C.c_int32), P]
    L.tgp_initpart_gpmetis.argtypes = L.tgp_initpart_block.argtypes
    L.tgp_initpart_gpmetis_dir.argtypes = L.tgp_initpart_block.argtypes
    L.tgp_release_device_memory.restype = None
    L.tgp_sort_moves.argtypes = [P, C.c_int64, C.c_int64, P, P, P]
    L.tgp_graph_upload.argtypes = [C.POINTER(CGraph), C.POINTER(CGraph)]
    L.tgp_graph_upload.restype = C.c_int
    L.tgp_graph_device_free.argtypes = [C.POINTER(CGraph)]
    L.tgp_graph_device_free.restype = None
    _lib = L
    return L


KERNEL_NAMES = ["match", "relabel", "contract", "pairs", "project", "gain", "token", "compact",
                "sort", "select", "commit", "cut", "scan", "misc", "match_chain"]


class TGPError(RuntimeError):
    def __init__(self, code, detail=""):
        msg = lib().tgp_strerror(code).decode()
        super().__init__(f"{msg} ({code})" + (f": {detail}" if detail else ""))
        self.code = code
        self.detail = detail


class Graph:
    """A CSR graph held in numpy arrays (int64 offsets, int32 ids/weights)."""

    def __init__(self, xadj, adjncy, adjwgt=None):
        self.xadj = np.ascontiguousarray(xadj, dtype=np.int64)
        self.adjncy = np.ascontiguousarray(adjncy, dtype=np.int32)
        if adjwgt is None:
            adjwgt = np.ones(len(self.adjncy), np.int32)
        self.adjwgt = np.ascontiguousarray(adjwgt, dtype=np.int32)
        self.n = len(self.xadj) - 1
        self.m = len(self.adjncy)

    def cgraph(self):
        return CGraph(self.n, self.m, self.xadj.ctypes.data, self.adjncy.ctypes.data,
                      self.adjwgt.ctypes.data)


def _take_cgraph(cg):
    n, m = cg.n, cg.m
    g = Graph(np.ctypeslib.as_array(C.cast(cg.xadj, C.POINTER(C.c_int64)), (n + 1,)).copy(),
              np.ctypeslib.as_array(C.cast(cg.adjncy, C.POINTER(C.c_int32)), (m,)).copy(),
              np.ctypeslib.as_array(C.cast(cg.adjwgt, C.POINTER(C.c_int32)), (m,)).copy())
    lib().tgp_graph_free(C.byref(cg))
    return g


def gen_rmat(scale, edge_factor, seed, weights="unit"):
    cg = CGraph()
    rc = lib().tgp_gen_rmat(scale, edge_factor, seed, 1 if weights == "hash16" else 0,
                            C.byref(cg))
    if rc != 0:
        raise TGPError(rc)
    return _take_cgraph(cg)


def gen_rmat_n(n, edge_factor, seed):
    """R-MAT (0.57, 0.19, 0.19, 0.05) on n vertices: the first edge_factor*n
    samples of the scale ceil(log2 n) model with both endpoints below n."""
    cg = CGraph()
    rc = lib().tgp_gen_rmat_n(n, edge_factor, seed, C.byref(cg))
    if rc != 0:
        raise TGPError(rc)
    return _take_cgraph(cg)


def gen_grid3d(nx, ny=None, nz=None):
    cg = CGraph()
    rc = lib().tgp_gen_grid3d(nx, ny or nx, nz or nx, C.byref(cg))
    if rc != 0:
        raise TGPError(rc)
    return _take_cgraph(cg)


def gen_er(n, mean_degree, seed):
    cg = CGraph()
    rc = lib().tgp_gen_er(n, mean_degree, seed, C.byref(cg))
    if rc != 0:
        raise TGPError(rc)
    return _take_cgraph(cg)


def read_graph_dir(path, n, m):
    cg = CGraph()
    rc = lib().tgp_read_graph_dir(path.encode(), n, m, C.byref(cg))
    if rc != 0:
        raise TGPError(rc, lib().tgp_last_error().decode())
    return _take_cgraph(cg)


def read_csr_files(row_path, col_path):
    """The README form's inputs: row.txt and column.txt, unit weights."""
    cg = CGraph()
    rc = lib().tgp_read_csr_files(os.fsencode(row_path), os.fsencode(col_path), C.byref(cg))
    if rc != 0:
        raise TGPError(rc, lib().tgp_last_error().decode())
    return _take_cgraph(cg)


def write_bcsr(path, g):
    """Binary CSR file ("TGPCSR01", n, m, xadj, adjncy, adjwgt)."""
    rc = lib().tgp_write_bcsr(os.fsencode(path), C.byref(g.cgraph()))
    if rc != 0:
        raise TGPError(rc, lib().tgp_last_error().decode())


def read_bcsr(path):
    cg = CGraph()
    rc = lib().tgp_read_bcsr(os.fsencode(path), C.byref(cg))
    if rc != 0:
        raise TGPError(rc, lib().tgp_last_error().decode())
    return _take_cgraph(cg)


def write_part_file(path, part):
    """METIS-style partition file: one part id per line, vertex order."""
    p = np.ascontiguousarray(part, dtype=np.int32)
    rc = lib().tgp_write_part_file(os.fsencode(path), p.ctypes.data, len(p))
    if rc != 0:
        raise TGPError(rc, lib().tgp_last_error().decode())


def convert_edges(path):
    """An edge-list file as conversion.py reads it (.mtx/.mmio, .tsv, .txt,
    .tsv_1), made simple, bidirected and free of isolated vertices."""
    cg = CGraph()
    rc = lib().tgp_convert_edges(os.fsencode(path), C.byref(cg))
    if rc != 0:
        raise TGPError(rc, f"cannot convert {path}")
    return _take_cgraph(cg)


def write_conversion_files(g, out_dir="."):
    """row.txt, column.txt, weight.txt, output_output.csr and temp.graph."""
    rc = lib().tgp_write_conversion_files(C.byref(g.cgraph()), os.fsencode(out_dir))
    if rc != 0:
        raise TGPError(rc, f"cannot write into {out_dir}")


def write_final_partition(path, part, k):
    part = np.ascontiguousarray(part, dtype=np.int32)
    rc = lib().tgp_write_final_partition(path.encode(), part.ctypes.data, len(part), k)
    if rc != 0:
        raise TGPError(rc)


def metis_text(g):
    cg = g.cgraph()
    buf = C.c_void_p()
    ln = C.c_int64()
    npairs = C.c_int64()
    rc = lib().tgp_metis_text(C.byref(cg), C.byref(buf), C.byref(ln), C.byref(npairs))
    if rc != 0:
        raise TGPError(rc)
    data = C.string_at(buf, ln.value)
    lib().tgp_free(buf)
    return data, npairs.value


def release_device_memory():
    """Returns the device blocks the library's pool keeps for reuse to the HIP
    runtime (tgp_release_device_memory)."""
    lib().tgp_release_device_memory()


class DeviceGraph:
    """A CSR copied once into device memory (HBM); free() returns it."""

    def __init__(self, g):
        self.n, self.m = g.n, g.m
        self._cg = CGraph()
        rc = lib().tgp_graph_upload(C.byref(g.cgraph()), C.byref(self._cg))
        if rc != 0:
            raise TGPError(rc)

    def pointers(self):
        return (self._cg.xadj, self._cg.adjncy, self._cg.adjwgt)

    def free(self):
        if self._cg.xadj:
            lib().tgp_graph_device_free(C.byref(self._cg))


class Stats:
    def __init__(self, cs, trace):
        self.levels = cs.levels
        self.passes = cs.passes
        self.coarsest_n = cs.coarsest_n
        self.coarsest_npairs = cs.coarsest_npairs
        self.match_rounds = cs.match_rounds
        self.seconds_total = cs.seconds_total
        self.seconds_initpart = cs.seconds_initpart
        self.fail_level = cs.fail_level
        self.fail_value = cs.fail_value
        self.sorts_skipped = cs.sorts_skipped
        self.level0_form = cs.level0_form
        self.kernel_ms = {KERNEL_NAMES[i]: cs.kernel_ms[i] for i in range(len(KERNEL_NAMES))}
        self.kernel_launches = {KERNEL_NAMES[i]: cs.kernel_launches[i]
                                for i in range(len(KERNEL_NAMES))}
        self.kernel_bytes = {KERNEL_NAMES[i]: cs.kernel_bytes[i]
                             for i in range(len(KERNEL_NAMES))}
        self.trace = trace


def partition(g, k, *, init=INIT_GPMETIS, init_part=None, wide=False, repair=False,
              coarsen_floor=0, profile=False, check_level=0, trace_cap=0, device_graph=None,
              pass_dump=None, learned=None, noalloc=False, balance_eps=0.0):
    """Partitions `g` (a Graph) into k parts on the GPU.

    Returns (part[n] int32, part_count[k] int32, edge_cut, Stats).  `init`
    chooses the initial partitioner of the coarsest graph: "gpmetis" (the
    reference seam, needs gpmetis on PATH), "block" (i*k/n_c), or a Python
    callable f(xadj, adjncy, adjwgt, npairs, k) -> part array.  `init_part`
    supplies fixed values instead (like a fixture part file).
    `device_graph`, if given, is a (xadj, adjncy, adjwgt) triple of device
    pointers holding the same CSR already resident in HBM.  `pass_dump` is a
    PASS_DUMP_FN called with each level's partition at every phase start and
    after every committed pass; `learned` a tgp_learned_model (learn.GNN.handle())
    whose learned passes run at the finest level before its refinement.
    `balance_eps` > 0 adds the balance stage at the finest level (every part
    at most floor((1 + eps) * ceil(n / k)) vertices; see tgp_options).
    `noalloc` selects tgp_partition_noalloc (same result; every device buffer
    from the library's fixed scratch; size limits TGP_NOALLOC_MAX_*).
    """
    L = lib()
    o = COptions()
    o.k = k
    o.wide = int(bool(wide))
    o.coarsen_floor = int(coarsen_floor)
    o.profile = int(bool(profile))
    o.check_level = int(check_level)
    o.repair = int(bool(repair))
    keep = []
    if init_part is not None:
        fixed = np.ascontiguousarray(init_part, dtype=np.int32)

        def init_cb(cgp, npairs, kk, out, user):
            nc = cgp.contents.n
            for i in range(nc):
                out[i] = int(fixed[i]) if i < len(fixed) else 0
            return 0
        cb = INITPART_FN(init_cb)
        keep.append(cb)
        o.initpart = C.cast(cb, C.c_void_p).value
    elif callable(init):
        def init_cb(cgp, npairs, kk, out, user):
            cgv = cgp.contents
            xa = np.ctypeslib.as_array(C.cast(cgv.xadj, C.POINTER(C.c_int64)), (cgv.n + 1,))
            ad = np.ctypeslib.as_array(C.cast(cgv.adjncy, C.POINTER(C.c_int32)), (max(cgv.m, 1),))
            wt = np.ctypeslib.as_array(C.cast(cgv.adjwgt, C.POINTER(C.c_int32)), (max(cgv.m, 1),))
            p = np.asarray(init(xa.copy(), ad[:cgv.m].copy(), wt[:cgv.m].copy(), npairs, kk),
                           dtype=np.int32)
            for i in range(cgv.n):
                out[i] = int(p[i])
            return 0
        cb = INITPART_FN(init_cb)
        keep.append(cb)
        o.initpart = C.cast(cb, C.c_void_p).value
    elif init == INIT_BLOCK:
        o.initpart = C.cast(L.tgp_initpart_block, C.c_void_p).value
    elif init == INIT_GPMETIS:
        o.initpart = C.cast(L.tgp_initpart_gpmetis, C.c_void_p).value
    else:
        raise ValueError(f"unknown initial partitioner {init!r}")
    if pass_dump is not None:
        o.pass_dump = C.cast(pass_dump, C.c_void_p).value
        keep.append(pass_dump)
    o.balance_eps = float(balance_eps)
    if learned is not None:
        o.learned = C.addressof(learned)
        keep.append(learned)
    if device_graph is not None:
        o.device_input = 1
        cg = CGraph(g.n, g.m, device_graph[0], device_graph[1], device_graph[2])
    else:
        cg = g.cgraph()
    part = np.zeros(g.n, np.int32)
    pc = np.zeros(k, np.int32)
    cut = C.c_int64()
    st = CStats()
    tbuf = None
    if trace_cap:
        tbuf = (CTrace * trace_cap)()
        st.trace = C.cast(tbuf, C.POINTER(CTrace))
        st.trace_cap = trace_cap
    entry = L.tgp_partition_noalloc if noalloc else L.tgp_partition
    rc = entry(C.byref(cg), C.byref(o), part.ctypes.data, pc.ctypes.data,
               C.addressof(cut), C.byref(st))
    trace = []
    if tbuf is not None:
        for i in range(min(st.trace_len, trace_cap)):
            t = tbuf[i]
            trace.append((t.kind, t.level, t.phase, t.pass_, t.a, t.b, t.c, t.d, t.hash))
    if rc != 0:
        err = TGPError(rc, L.tgp_last_error().decode())
        err.fail_level = st.fail_level
        err.fail_value = st.fail_value
        err.trace = trace
        raise err
    return part, pc, cut.value, Stats(st, trace)


def balance(g, part, k, eps, *, wide=False, device_graph=None, trace_cap=0):
    """The balance stage alone (tgp_balance) on partition `part` of g: returns
    (part, part_count, edge_cut, Stats) with every part within
    floor((1 + eps) * ceil(n / k)) vertices."""
    L = lib()
    o = COptions()
    o.k = k
    o.wide = int(bool(wide))
    o.balance_eps = float(eps)
    if device_graph is not None:
        o.device_input = 1
        cg = CGraph(g.n, g.m, device_graph[0], device_graph[1], device_graph[2])
    else:
        cg = g.cgraph()
    p = np.array(part, dtype=np.int32)
    pc = np.zeros(k, np.int32)
    cut = C.c_int64()
    st = CStats()
    tbuf = None
    if trace_cap:
        tbuf = (CTrace * trace_cap)()
        st.trace = C.cast(tbuf, C.POINTER(CTrace))
        st.trace_cap = trace_cap
    rc = L.tgp_balance(C.byref(cg), C.byref(o), p.ctypes.data, pc.ctypes.data,
                       C.addressof(cut), C.byref(st))
    trace = []
    if tbuf is not None:
        for i in range(min(st.trace_len, trace_cap)):
            t = tbuf[i]
            trace.append((t.kind, t.level, t.phase, t.pass_, t.a, t.b, t.c, t.d, t.hash))
    if rc != 0:
        raise TGPError(rc, L.tgp_last_error().decode())
    return p, pc, cut.value, Stats(st, trace)


def _batch_items(L, graphs, ks, init, wide, repair, coarsen_floor, device_graphs, profile,
                 workdir, balance_eps):
    """tgp_batch_item array for graphs/ks, their output arrays, and the
    objects the items point into (kept alive by the caller)."""
    cnt = len(graphs)
    if len(ks) != cnt:
        raise ValueError("one k per graph")
    dirs = None
    if init == INIT_BLOCK:
        fn = C.cast(L.tgp_initpart_block, C.c_void_p).value
    elif init == INIT_GPMETIS and workdir is not None:
        fn = C.cast(L.tgp_initpart_gpmetis_dir, C.c_void_p).value
        dirs = []
        for i in range(cnt):
            d = os.path.join(os.path.abspath(workdir), f"graph{i}")
            os.makedirs(d, exist_ok=True)
            dirs.append(CGpmetisDir(d.encode(), 0, 0))
    elif init == INIT_GPMETIS:
        fn = C.cast(L.tgp_initpart_gpmetis, C.c_void_p).value
    else:
        raise ValueError(f"batched form takes init 'block' or 'gpmetis', not {init!r}")
    items = (CBatchItem * max(cnt, 1))()
    outs = []
    for i, g in enumerate(graphs):
        it = items[i]
        if dirs is not None:
            it.o.initpart_user = C.addressof(dirs[i])
        if device_graphs is not None:
            d = device_graphs[i]
            it.g = CGraph(g.n, g.m, d[0], d[1], d[2])
            it.o.device_input = 1
        else:
            it.g = g.cgraph()
        it.o.k = int(ks[i])
        it.o.wide = int(bool(wide))
        it.o.repair = int(bool(repair))
        it.o.coarsen_floor = int(coarsen_floor)
        it.o.profile = int(bool(profile))
        it.o.initpart = fn
        it.o.balance_eps = float(balance_eps)
        part = np.zeros(g.n, np.int32)
        pc = np.zeros(int(ks[i]), np.int32)
        st = CStats()
        it.part = part.ctypes.data
        it.part_count = pc.ctypes.data
        it.stats = C.pointer(st)
        outs.append((part, pc, st))
    return items, outs, dirs


def batch_tuning(mean_degree):
    """(lanes, heavy_threshold, sort_tile) the batched engine uses for a graph
    of this mean degree (arcs / vertices), from its tuning table
    (tgp_batch_tuning)."""
    L = lib()
    a, h, t = C.c_int32(), C.c_int32(), C.c_int32()
    rc = L.tgp_batch_tuning(float(mean_degree), C.byref(a), C.byref(h), C.byref(t))
    if rc != 0:
        raise TGPError(rc, L.tgp_last_error().decode())
    return a.value, h.value, t.value


def batch_autotune(graphs, ks, *, path=None, device_graphs=None, wide=False, coarsen_floor=0):
    """Measures the batched engine's settings on these graphs (a sample of the
    workload; repaired coarsening, block initial partition) and writes the
    tuning table to `path` if given (tgp_batch_autotune).  Returns (report,
    results): the measurements and the table as text, and per graph (part,
    part_count, edge_cut, rc) of the chosen setting."""
    L = lib()
    items, outs, _keep = _batch_items(L, graphs, ks, INIT_BLOCK, wide, True, coarsen_floor,
                                      device_graphs, False, None, 0.0)
    rep = C.c_void_p()
    rc = L.tgp_batch_autotune(items, len(graphs), path.encode() if path else None,
                              C.byref(rep))
    text = ""
    if rep.value:
        text = C.string_at(rep.value).decode()
        L.tgp_free(rep)
    if rc != 0:
        raise TGPError(rc, L.tgp_last_error().decode() + "\n" + text)
    return text, [(part, pc, items[i].edge_cut, items[i].rc)
                  for i, (part, pc, _st) in enumerate(outs)]


def partition_batch(graphs, ks, *, init=INIT_BLOCK, wide=False, repair=False, coarsen_floor=0,
                    workers=8, device_graphs=None, profile=False, workdir=None,
                    balance_eps=0.0):
    """Partitions many independent graphs through tgp_partition_batch.

    Each result equals what partition() returns for that graph alone (bit for
    bit); up to `workers` graphs run at once, each on its own HIP stream.
    `init` is "block" or "gpmetis"; with "gpmetis" and a `workdir`, graph i
    calls gpmetis in its own directory <workdir>/graph<i> (created), so the
    calls run concurrently on host threads; without one they share the
    working directory and run one at a time.  `device_graphs`, if given,
    holds one (xadj, adjncy, adjwgt) triple of device pointers per graph.
    Returns a list of (part, part_count, edge_cut, Stats); raises TGPError
    naming the first failing graph.
    """
    L = lib()
    cnt = len(graphs)
    items, outs, _keep = _batch_items(L, graphs, ks, init, wide, repair, coarsen_floor,
                                      device_graphs, profile, workdir, balance_eps)
    rc = L.tgp_partition_batch(items, cnt, int(workers))
    if rc != 0:
        bad = next(i for i in range(cnt) if items[i].rc != 0)
        raise TGPError(rc, f"graph {bad}: {L.tgp_strerror(items[bad].rc).decode()}")
    return [(part, pc, items[i].edge_cut, Stats(st, [])) for i, (part, pc, st) in enumerate(outs)]
