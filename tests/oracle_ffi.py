"""ctypes binding of the CPU oracle (oracle/libtgp_oracle.so).

Test infrastructure: only tests, the smoke check and bench.py's CPU-baseline
leg import this module, and only to check or time against the oracle.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libtgp_oracle.so")

INIT_BLOCK, INIT_ARRAY, INIT_GPMETIS = 0, 1, 2
TR_LEVEL, TR_INIT, TR_PASS, TR_REFINE, TR_FINAL = 1, 2, 3, 4, 5


class Options(C.Structure):
    _fields_ = [("k", C.c_int32), ("wide", C.c_int32), ("coarsen_floor", C.c_int64),
                ("init_mode", C.c_int32), ("repair", C.c_int32),
                ("init_part", C.POINTER(C.c_int32)), ("init_part_len", C.c_int64)]


class Trace(C.Structure):
    _fields_ = [("kind", C.c_int32), ("level", C.c_int32), ("phase", C.c_int32),
                ("pass_", C.c_int32), ("a", C.c_int64), ("b", C.c_int64),
                ("c", C.c_int64), ("d", C.c_int64), ("hash", C.c_uint64)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i64, i32 = C.c_int64, C.c_int32
        L.tgpo_partition.argtypes = [i64, P, P, P, C.POINTER(Options), P, P, P,
                                     C.POINTER(Trace), i64, P]
        L.tgpo_partition.restype = C.c_int
        L.tgpo_match.argtypes = [i64, P, P, P, P]
        L.tgpo_relabel.argtypes = [i64, P, P, P]
        L.tgpo_relabel.restype = i64
        L.tgpo_contract.argtypes = [i64, P, P, P, P, P, i64, P, P, P]
        L.tgpo_contract.restype = i64
        L.tgpo_contract_repaired.argtypes = [i64, P, P, P, P, P, i64, P, P, P]
        L.tgpo_contract_repaired.restype = i64
        L.tgpo_metis_text.argtypes = [i64, P, P, P, C.c_char_p, i64, P]
        L.tgpo_metis_text.restype = i64
        L.tgpo_edge_cut.argtypes = [i64, P, P, P, i64, P, i32, i32]
        L.tgpo_edge_cut.restype = i64
        L.tgpo_gains.argtypes = [i64, P, P, P, i64, P, P, i32, i32, P, P]
        L.tgpo_bitonic.argtypes = [P, P, i64]
        L.tgpo_refine.argtypes = [i64, P, P, P, i64, i32, i32, P, P, P,
                                  C.POINTER(Trace), i64, P]
        L.tgpo_final_text.argtypes = [i64, P, i32, C.c_char_p, i64]
        L.tgpo_final_text.restype = i64
        L.tgpo_plant_mistake.argtypes = [i32]
        L.tgpo_set_threads.argtypes = [i32]
        L.tgpo_set_threads.restype = i32
        L.tgpo_hash.argtypes = [P, i64]
        L.tgpo_hash.restype = C.c_uint64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def csr_arrays(xadj, adj, w=None):
    xadj = np.ascontiguousarray(xadj, dtype=np.int64)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    w = np.ones(len(adj), np.int32) if w is None else np.ascontiguousarray(w, dtype=np.int32)
    return xadj, adj, w


def match(xadj, adj, w=None):
    xadj, adj, w = csr_arrays(xadj, adj, w)
    n = len(xadj) - 1
    m = np.empty(n, np.int32)
    lib().tgpo_match(n, _p(xadj), _p(adj), _p(w), _p(m))
    return m


def relabel(match_arr):
    match_arr = np.ascontiguousarray(match_arr, dtype=np.int32)
    n = len(match_arr)
    new_ids = np.empty(n, np.int32)
    tb = np.full(2 * n, -1, np.int32)
    n_c = lib().tgpo_relabel(n, _p(match_arr), _p(new_ids), _p(tb))
    return new_ids, tb[:2 * n_c], n_c


def contract(xadj, adj, w, match_arr, new_ids, n_c):
    xadj, adj, w = csr_arrays(xadj, adj, w)
    n = len(xadj) - 1
    match_arr = np.ascontiguousarray(match_arr, dtype=np.int32)
    new_ids = np.ascontiguousarray(new_ids, dtype=np.int32)
    cap = max(1, 2 * int(xadj[n]))
    ox = np.zeros(n_c + 1, np.int64)
    oa = np.zeros(cap, np.int32)
    ow = np.zeros(cap, np.int32)
    m_c = lib().tgpo_contract(n, _p(xadj), _p(adj), _p(w), _p(match_arr), _p(new_ids),
                              n_c, _p(ox), _p(oa), _p(ow))
    return ox, oa[:m_c].copy(), ow[:m_c].copy()


def contract_repaired(xadj, adj, w, tb, new_ids, n_c):
    xadj, adj, w = csr_arrays(xadj, adj, w)
    n = len(xadj) - 1
    tb = np.ascontiguousarray(tb, dtype=np.int32)
    new_ids = np.ascontiguousarray(new_ids, dtype=np.int32)
    cap = max(1, int(xadj[n]))
    ox = np.zeros(n_c + 1, np.int64)
    oa = np.zeros(cap, np.int32)
    ow = np.zeros(cap, np.int32)
    m_c = lib().tgpo_contract_repaired(n, _p(xadj), _p(adj), _p(w), _p(tb), _p(new_ids),
                                       n_c, _p(ox), _p(oa), _p(ow))
    return ox, oa[:m_c].copy(), ow[:m_c].copy()


def metis_text(xadj, adj, w=None):
    xadj, adj, w = csr_arrays(xadj, adj, w)
    n = len(xadj) - 1
    npairs = np.zeros(1, np.int64)
    size = lib().tgpo_metis_text(n, _p(xadj), _p(adj), _p(w), None, 0, _p(npairs))
    buf = C.create_string_buffer(int(size) + 1)
    lib().tgpo_metis_text(n, _p(xadj), _p(adj), _p(w), buf, size, _p(npairs))
    return buf.raw[:size], int(npairs[0])


def edge_cut(xadj, adj, w, nedges, part, k, wide=0):
    xadj, adj, w = csr_arrays(xadj, adj, w)
    part = np.ascontiguousarray(part, dtype=np.int32)
    return int(lib().tgpo_edge_cut(len(xadj) - 1, _p(xadj), _p(adj), _p(w), nedges,
                                   _p(part), k, wide))


def gains(xadj, adj, w, nedges, part, pc, k, phase):
    xadj, adj, w = csr_arrays(xadj, adj, w)
    n = len(xadj) - 1
    part = np.ascontiguousarray(part, dtype=np.int32)
    pc = np.ascontiguousarray(pc, dtype=np.int32)
    g = np.zeros(n, np.float32)
    t = np.zeros(n, np.int32)
    lib().tgpo_gains(n, _p(xadj), _p(adj), _p(w), nedges, _p(part), _p(pc), k, phase,
                     _p(g), _p(t))
    return g, t


def bitonic(gain, ids):
    gain = np.array(gain, dtype=np.float32)
    ids = np.array(ids, dtype=np.int32)
    lib().tgpo_bitonic(_p(gain), _p(ids), len(gain))
    return gain, ids


def _trace_list(buf, count):
    out = []
    for i in range(min(count, len(buf))):
        t = buf[i]
        out.append((t.kind, t.level, t.phase, t.pass_, t.a, t.b, t.c, t.d, t.hash))
    return out


def refine(xadj, adj, w, nedges, k, part, pc, cut, wide=0, trace_cap=4096):
    xadj, adj, w = csr_arrays(xadj, adj, w)
    n = len(xadj) - 1
    part = np.array(part, dtype=np.int32)
    pc = np.array(pc, dtype=np.int32)
    cutv = np.array([cut], np.int64)
    tr = (Trace * trace_cap)()
    tl = np.zeros(1, np.int64)
    lib().tgpo_refine(n, _p(xadj), _p(adj), _p(w), nedges, k, wide, _p(part), _p(pc),
                      _p(cutv), tr, trace_cap, _p(tl))
    return part, pc, int(cutv[0]), _trace_list(tr, int(tl[0]))


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"oracle error {code}")
        self.code = code


def partition(xadj, adj, w, k, wide=0, coarsen_floor=0, init_mode=INIT_BLOCK,
              init_part=None, trace_cap=1 << 16, repair=0):
    xadj, adj, w = csr_arrays(xadj, adj, w)
    n = len(xadj) - 1
    o = Options()
    o.k, o.wide, o.coarsen_floor, o.init_mode = k, wide, coarsen_floor, init_mode
    o.repair = repair
    ip = None
    if init_part is not None:
        ip = np.ascontiguousarray(init_part, dtype=np.int32)
        o.init_part = ip.ctypes.data_as(C.POINTER(C.c_int32))
        o.init_part_len = len(ip)
    part = np.zeros(n, np.int32)
    pc = np.zeros(k, np.int32)
    cut = np.zeros(1, np.int64)
    tr = (Trace * trace_cap)()
    tl = np.zeros(1, np.int64)
    rc = lib().tgpo_partition(n, _p(xadj), _p(adj), _p(w), C.byref(o), _p(part), _p(pc),
                              _p(cut), tr, trace_cap, _p(tl))
    trace = _trace_list(tr, int(tl[0]))
    if rc != 0:
        err = OracleError(rc)
        err.trace = trace
        raise err
    return part, pc, int(cut[0]), trace


def final_text(part, k):
    part = np.ascontiguousarray(part, dtype=np.int32)
    size = lib().tgpo_final_text(len(part), _p(part), k, None, 0)
    buf = C.create_string_buffer(int(size) + 1)
    lib().tgpo_final_text(len(part), _p(part), k, buf, size)
    return buf.raw[:size]


class planted:
    """Context manager: the oracle runs with planted mistake `which`."""
    def __init__(self, which):
        self.which = which

    def __enter__(self):
        lib().tgpo_plant_mistake(self.which)

    def __exit__(self, *exc):
        lib().tgpo_plant_mistake(0)


def set_threads(threads):
    """Threads of the oracle's parallel loops (0: keep); returns the count in use."""
    return int(lib().tgpo_set_threads(threads))


def hash_part(part):
    part = np.ascontiguousarray(part, dtype=np.int32)
    return int(lib().tgpo_hash(_p(part), len(part)))
