"""CPU-side checks of the native library (no GPU work is started here).

* libtgp.so loads and exports every function include/tgp.h declares;
* the host halves of the reference's file contract (METIS text, the
  finalpartition.txt writer, the row/column/weight readers) match the oracle
  byte for byte;
* the `gpu_part` command reproduces the reference's argument errors;
* the workload generators are deterministic and produce what
  preprocess/conversion.py would: symmetric, no self loops, no isolated
  vertices, ascending rows.
"""
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

import graphs
import thegraphpartitioning_amd as tgp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "tgp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(tgp_[a-z0-9_]+)\s*\(", text))
    types = set(re.findall(r"\(\*\s*(tgp_[a-z0-9_]+)\)", text))
    return sorted(names - types)


def test_library_exports_every_declared_symbol():
    lib = tgp.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_oracle_library_exports():
    import oracle_ffi
    text = open(os.path.join(ROOT, "oracle", "tgp_oracle.h")).read()
    for n in re.findall(r"\b(tgpo_[a-z_]+)\s*\(", text):
        assert hasattr(oracle_ffi.lib(), n), n


@pytest.mark.parametrize("seed", range(5))
def test_metis_text_matches_oracle(oracle, seed):
    x, a, w = graphs.random_graph(300, 1200, seed, sym=seed % 2 == 0, wmax=1 + seed,
                                  dup=seed == 4, selfloops=seed == 4)
    mine, np1 = tgp.metis_text(tgp.Graph(x, a, w))
    ref, np2 = oracle.metis_text(x, a, w)
    assert mine == ref and np1 == np2


def test_final_partition_writer_matches_oracle(oracle, tmp_path):
    rng = np.random.default_rng(3)
    part = rng.integers(-1, 9, size=500).astype(np.int32)
    path = str(tmp_path / "finalpartition.txt")
    tgp.write_final_partition(path, part, 7)
    assert open(path, "rb").read() == oracle.final_text(part, 7)


def write_dir(d, xadj, adj, w):
    os.makedirs(d, exist_ok=True)
    np.savetxt(os.path.join(d, "row.txt"), xadj, fmt="%d")
    np.savetxt(os.path.join(d, "column.txt"), adj, fmt="%d")
    np.savetxt(os.path.join(d, "weight.txt"), w, fmt="%d")


def test_read_graph_dir_roundtrip(tmp_path):
    x, a, w = graphs.random_graph(200, 700, 1, wmax=5)
    d = str(tmp_path / "g")
    write_dir(d, x, a, w)
    g = tgp.read_graph_dir(d, len(x) - 1, len(a))
    np.testing.assert_array_equal(g.xadj, x)
    np.testing.assert_array_equal(g.adjncy, a)
    np.testing.assert_array_equal(g.adjwgt, w)
    with pytest.raises(tgp.TGPError):
        tgp.read_graph_dir(d, len(x) - 1, len(a) + 5)          # short column.txt
    with pytest.raises(tgp.TGPError):
        tgp.read_graph_dir(str(tmp_path / "missing"), 3, 3)


def run_cli(args, cwd):
    return subprocess.run([tgp.CLI_PATH] + args, cwd=cwd, stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=60)


def test_cli_usage_and_argument_errors(tmp_path):
    r = run_cli(["a", "b"], str(tmp_path))
    assert r.returncode == 1
    assert "Usage:" in r.stderr and "<directory> <nvtxs> <nedges> <num_partitions>" in r.stderr
    r = run_cli(["d", "0", "5", "2"], str(tmp_path))
    assert r.returncode == 1
    assert r.stderr.strip() == "Invalid input: nvtxs, nedges, and numPartitions must be positive."
    r = run_cli([str(tmp_path / "nope"), "4", "6", "2"], str(tmp_path))
    assert r.returncode == 1 and "Error opening file" in r.stderr
    d = str(tmp_path / "g")
    x, a, w = graphs.path_graph(4)
    write_dir(d, x, a, w)
    r = run_cli([d, "4", "9", "2"], str(tmp_path))
    assert r.returncode == 1 and "Error reading file" in r.stderr


def test_generators_are_canonical_and_deterministic():
    g1 = tgp.gen_rmat(12, 8, 12345)
    g2 = tgp.gen_rmat(12, 8, 12345)
    np.testing.assert_array_equal(g1.adjncy, g2.adjncy)
    for g in (g1, tgp.gen_grid3d(6, 5, 4), tgp.gen_er(3000, 6.0, 2)):
        x, a = g.xadj, g.adjncy
        deg = np.diff(x)
        assert (deg > 0).all()                                  # no isolated vertices
        src = np.repeat(np.arange(g.n), deg)
        assert not (src == a).any()                             # no self loops
        fwd = set(zip(src.tolist(), a.tolist()))
        assert all((v, u) in fwd for u, v in fwd)               # symmetric
        for v in range(0, g.n, max(1, g.n // 50)):
            row = a[x[v]:x[v + 1]]
            assert (np.diff(row) > 0).all()                     # ascending, deduplicated
    grid = tgp.gen_grid3d(6, 5, 4)
    assert grid.n == 120 and grid.m == 2 * (5 * 5 * 4 + 6 * 4 * 4 + 6 * 5 * 3)


def test_rmat_fixture_digest():
    """Pins the generator output so workloads stay identical across changes."""
    g = tgp.gen_rmat(10, 4, 0x7C0DE5EED)
    h = hashlib.sha256(g.xadj.tobytes() + g.adjncy.tobytes() + g.adjwgt.tobytes()).hexdigest()
    golden = open(os.path.join(ROOT, "tests", "golden", "rmat10_ef4.sha256")).read().split()[0]
    assert h == golden


def test_cli_other_input_forms_errors(tmp_path):
    """README form (row file, column file, k) and binary form (graph.bcsr, k):
    argument and file errors are reported before any device work."""
    row = tmp_path / "row.txt"
    row.write_text("0\n1\n2\n")
    r = run_cli([str(row), str(tmp_path / "nope.txt"), "2"], str(tmp_path))
    assert r.returncode == 1 and "Error opening file" in r.stderr
    r = run_cli([str(row), str(row), "0"], str(tmp_path))
    assert r.returncode == 1
    assert r.stderr.strip() == "Invalid input: nvtxs, nedges, and numPartitions must be positive."
    r = run_cli([str(tmp_path / "g.bcsr"), "2"], str(tmp_path))
    assert r.returncode == 1 and "Error opening file" in r.stderr
    r = run_cli([str(row), str(row), "2", "--bogus"], str(tmp_path))
    assert r.returncode == 1 and "Unknown option: --bogus" in r.stderr



def test_gpmetis_seam_in_a_directory(tmp_path, monkeypatch, stub_gpmetis):
    """tgp_initpart_gpmetis_dir writes the same METIS text as the working-
    directory seam (itself checked against the oracle above), runs gpmetis on
    the file in its own directory and reads the part file by the same rules."""
    import ctypes as C
    log = stub_gpmetis
    L = tgp.lib()
    xadj, adj, w = graphs.random_graph(60, 150, 4, wmax=5)
    g = tgp.Graph(xadj, adj, w)
    cg = g.cgraph()
    k = 4
    cwd = tmp_path / "cwd"
    cwd.mkdir()
    monkeypatch.chdir(cwd)
    status = C.c_int(-1)
    part_cwd = (C.c_int32 * g.n)()
    assert L.tgp_initpart_gpmetis(C.byref(cg), 0, k, part_cwd,
                                  C.cast(C.byref(status), C.c_void_p)) == 0
    assert status.value == 0
    own = tmp_path / "graph 0"        # a space in the name: the path is quoted
    own.mkdir()
    d = tgp.CGpmetisDir(str(own).encode(), -1, 0)
    part_dir = (C.c_int32 * g.n)()
    assert L.tgp_initpart_gpmetis_dir(C.byref(cg), 0, k, part_dir,
                                      C.cast(C.byref(d), C.c_void_p)) == 0
    assert d.status == 0
    assert (cwd / "output.graph").read_bytes() == tgp.metis_text(g)[0]
    assert (own / "output.graph").read_bytes() == tgp.metis_text(g)[0]
    expect = [(3 * i + 1) % k for i in range(g.n)]
    assert list(part_cwd) == expect and list(part_dir) == expect
    calls = log.read_text().splitlines()
    assert calls == [f"./output.graph {k}", f"{own}/output.graph {k}"]
    for name in (str(tmp_path / "it's"), ""):
        bad = tgp.CGpmetisDir(name.encode(), 0, 0)
        assert L.tgp_initpart_gpmetis_dir(C.byref(cg), 0, k, part_dir,
                                          C.cast(C.byref(bad), C.c_void_p)) == -1


def test_sort_moves_rejects_bad_arguments():
    """tgp_sort_moves checks its sizes and pointers before touching the device."""
    import ctypes as C
    L = tgp.lib()
    buf = (C.c_float * 4)()
    ids = (C.c_int32 * 4)()
    assert L.tgp_sort_moves(None, 2, 4, C.cast(buf, C.c_void_p), C.cast(ids, C.c_void_p),
                            None) == -1
    assert L.tgp_sort_moves(C.cast(buf, C.c_void_p), 5, 4, C.cast(buf, C.c_void_p),
                            C.cast(ids, C.c_void_p), None) == -1          # M > len
    assert L.tgp_sort_moves(C.cast(buf, C.c_void_p), -1, 4, C.cast(buf, C.c_void_p),
                            C.cast(ids, C.c_void_p), None) == -1
    assert L.tgp_sort_moves(C.cast(buf, C.c_void_p), 0, 0, C.cast(buf, C.c_void_p),
                            C.cast(ids, C.c_void_p), None) == 0           # nothing to sort


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says so,
    before touching a device."""
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
