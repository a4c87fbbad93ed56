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
import ctypes as C
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


STUB_DIR = os.path.join(ROOT, "tests", "golden", "stub_bin")


def seam_call(tmp_path, monkeypatch, mode, k=4):
    """Runs tgp_initpart_gpmetis (the working-directory seam) and the
    directory seam with tests/golden/stub_bin/gpmetis in TGP_STUB_MODE `mode`;
    returns (rc, part, status, last_error) of each."""
    import ctypes as C
    monkeypatch.setenv("PATH", f"{STUB_DIR}:{os.environ['PATH']}")
    monkeypatch.setenv("TGP_STUB_MODE", mode)
    L = tgp.lib()
    xadj, adj, w = graphs.random_graph(40, 100, 9, wmax=3)
    g = tgp.Graph(xadj, adj, w)
    cg = g.cgraph()
    cwd = tmp_path / f"cwd_{mode}"
    cwd.mkdir()
    monkeypatch.chdir(cwd)
    out = []
    status = C.c_int(-1)
    part = (C.c_int32 * g.n)(*([77] * g.n))
    rc = L.tgp_initpart_gpmetis(C.byref(cg), 0, k, part, C.cast(C.byref(status), C.c_void_p))
    out.append((rc, list(part), status.value, L.tgp_last_error().decode()))
    own = tmp_path / f"own_{mode}"
    own.mkdir()
    d = tgp.CGpmetisDir(str(own).encode(), -1, 0)
    part = (C.c_int32 * g.n)(*([77] * g.n))
    rc = L.tgp_initpart_gpmetis_dir(C.byref(cg), 0, k, part, C.cast(C.byref(d), C.c_void_p))
    out.append((rc, list(part), d.status, L.tgp_last_error().decode()))
    return g, cwd, own, out


def test_gpmetis_seam_fault_modes(tmp_path, monkeypatch):
    """The reference's handling of its METIS seam (gpu_part.cu:855-863,
    1162-1173), with the stand-in's injected faults:
      ok      parts i*k/n read back;
      fail    gpmetis exits 3: "Error: gpmetis failed with return code 768"
              (system()'s status), TGP_EINITPART;
      nofile  no part file: "Error: Cannot open <graph>.part.<k>", TGP_EPARTFILE;
      short   half the lines: the missing reads leave 0;
      range   every value k+5: out-of-range ids become 0."""
    k = 4
    g, cwd, own, out = seam_call(tmp_path, monkeypatch, "", k)
    n = g.n
    expect = [i * k // n for i in range(n)]
    for rc, part, status, _ in out:
        assert rc == 0 and status == 0 and part == expect
    assert (cwd / "output.graph").read_bytes() == tgp.metis_text(g)[0]

    _, _, own, out = seam_call(tmp_path, monkeypatch, "fail", k)
    for rc, part, status, err in out:
        assert rc == -4 and status == 768
        assert err == "Error: gpmetis failed with return code 768"

    _, _, own, out = seam_call(tmp_path, monkeypatch, "nofile", k)
    assert out[0][0] == -5 and out[0][3] == f"Error: Cannot open output.graph.part.{k}"
    assert out[1][0] == -5 and out[1][3] == f"Error: Cannot open {own}/output.graph.part.{k}"

    _, _, _, out = seam_call(tmp_path, monkeypatch, "short", k)
    half = [i * k // n for i in range(n // 2)] + [0] * (n - n // 2)
    for rc, part, _, _ in out:
        assert rc == 0 and part == half

    _, _, _, out = seam_call(tmp_path, monkeypatch, "range", k)
    for rc, part, _, _ in out:
        assert rc == 0 and part == [0] * n


def test_gpmetis_seam_matches_oracle_part_file_rules(tmp_path, monkeypatch, oracle):
    """The oracle's gpmetis mode reads the same stand-in outputs the same way:
    the partition it starts from (visible in its INIT record's cut) equals
    the library seam's, for the valid and the sanitised part files."""
    x, a, w = graphs.path_graph(12)
    monkeypatch.setenv("PATH", f"{STUB_DIR}:{os.environ['PATH']}")
    for mode in ("", "short", "range"):
        monkeypatch.setenv("TGP_STUB_MODE", mode)
        d = tmp_path / f"o_{mode}"
        d.mkdir()
        monkeypatch.chdir(d)
        _, _, _, tr = oracle.partition(x, a, w, 3, init_mode=oracle.INIT_GPMETIS)
        init = [t for t in tr if t[0] == oracle.TR_INIT][0]
        lines = open(d / "output.graph.part.3").read().split()
        vals = [int(v) for v in lines] + [0] * (12 - len(lines))
        vals = [v if 0 <= v < 3 else 0 for v in vals]
        assert init[6] == oracle.edge_cut(x, a, w, init[5], vals, 3)
    monkeypatch.setenv("TGP_STUB_MODE", "nofile")
    d = tmp_path / "o_nofile"
    d.mkdir()
    monkeypatch.chdir(d)
    with pytest.raises(oracle.OracleError) as ei:
        oracle.partition(x, a, w, 3, init_mode=oracle.INIT_GPMETIS)
    assert ei.value.code == -5
    monkeypatch.setenv("TGP_STUB_MODE", "fail")
    with pytest.raises(oracle.OracleError) as ei:
        oracle.partition(x, a, w, 3, init_mode=oracle.INIT_GPMETIS)
    assert ei.value.code == -4


def test_row_offsets_are_read_as_int64(tmp_path):
    """row.txt values past INT_MAX are kept exactly (int64 offsets), so the
    text forms can describe more than 2^31 arcs in wide mode instead of
    wrapping into a different, possibly valid-looking CSR."""
    (tmp_path / "row.txt").write_text("0\n1\n4294967297\n")
    (tmp_path / "column.txt").write_text("1\n0\n")
    (tmp_path / "weight.txt").write_text("1\n1\n")
    g = tgp.read_csr_files(str(tmp_path / "row.txt"), str(tmp_path / "column.txt"))
    assert g.xadj.tolist() == [0, 1, 4294967297]
    g = tgp.read_graph_dir(str(tmp_path), 2, 2)
    assert g.xadj.tolist() == [0, 1, 4294967297]


def test_ctypes_mirrors_match_the_c_structs(tmp_path):
    """The Python mirrors of include/tgp.h's structs have the C layout: a
    program built by the host compiler prints each struct's size and every
    field's offset, which must equal the ctypes ones."""
    mirrors = {"tgp_graph": tgp.CGraph, "tgp_options": tgp.COptions,
               "tgp_gpmetis_dir": tgp.CGpmetisDir, "tgp_trace": tgp.CTrace,
               "tgp_stats": tgp.CStats, "tgp_batch_item": tgp.CBatchItem}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tgp.h"', 'int main(void) {']
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            cf = fname.rstrip("_")  # "pass_" mirrors the C field "pass"
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {cf}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True, timeout=60)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True,
                         timeout=30).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
