"""The gpu_part command's three input forms give the same partition.

Reference form `<dir> <n> <m> <k>`, the README form `<row.txt> <column.txt>
<k>` (unit weights) and the binary CSR form `<graph.bcsr> <k>` on the same
unit-weight graph, each writing ./finalpartition.txt; the METIS-style part
file written with --part-out lists the same parts.
"""
import os
import subprocess

import numpy as np
import pytest

import thegraphpartitioning_amd as tgp

pytestmark = pytest.mark.gpu


def run(args, cwd):
    r = subprocess.run([tgp.CLI_PATH] + args, cwd=cwd, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return open(os.path.join(cwd, "finalpartition.txt")).read()


def test_cli_input_forms_agree(tmp_path):
    g = tgp.gen_rmat(11, 8, 21)
    d = tmp_path / "g"
    d.mkdir()
    np.savetxt(d / "row.txt", g.xadj, fmt="%d")
    np.savetxt(d / "column.txt", g.adjncy, fmt="%d")
    np.savetxt(d / "weight.txt", g.adjwgt, fmt="%d")
    opts = ["--init=block", "--repair"]
    for name in ("a", "b", "c"):
        (tmp_path / name).mkdir()
    bcsr = str(tmp_path / "g.bcsr")
    part_path = str(tmp_path / "g.part.4")
    a = run([str(d), str(g.n), str(g.m), "4"] + opts +
            [f"--part-out={part_path}", f"--write-bcsr={bcsr}"], str(tmp_path / "a"))
    b = run([str(d / "row.txt"), str(d / "column.txt"), "4"] + opts, str(tmp_path / "b"))
    c = run([bcsr, "4"] + opts, str(tmp_path / "c"))
    assert a == b == c
    part, _, _, _ = tgp.partition(g, 4, init=tgp.INIT_BLOCK, repair=True)
    assert np.loadtxt(part_path, dtype=np.int64).tolist() == part.tolist()
    ref = tmp_path / "ref.txt"
    tgp.write_final_partition(str(ref), part, 4)
    assert a == ref.read_text()
