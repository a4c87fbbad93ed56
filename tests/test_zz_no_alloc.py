"""tgp_partition_noalloc against tgp_partition, through the C ABI.

The entry point that allocates nothing must give the same partition, part
counts, cut and trace as tgp_partition on the same input, on consecutive
calls that share its device scratch (the second call with a smaller graph, so
stale contents of the first would show), and reject graphs beyond its limits
before doing any work.
"""
import numpy as np
import pytest

import thegraphpartitioning_amd as tgp

pytestmark = pytest.mark.gpu


def both(g, k, **kw):
    a = tgp.partition(g, k, check_level=1, trace_cap=1 << 14, **kw)
    b = tgp.partition(g, k, check_level=1, trace_cap=1 << 14, noalloc=True, **kw)
    return a, b


def assert_same(a, b, label):
    assert a[2] == b[2], f"{label}: cut {a[2]} vs {b[2]}"
    np.testing.assert_array_equal(a[0], b[0], err_msg=label)
    np.testing.assert_array_equal(a[1], b[1], err_msg=label)
    assert [tuple(t) for t in a[3].trace] == [tuple(t) for t in b[3].trace], label


def test_noalloc_equals_partition_on_consecutive_calls():
    # first call: near the size limits (8.0 M arcs, hub rows of 25 K arcs)
    big = tgp.gen_rmat(19, 8, 0x7C0DE5EED + 19)
    print(f"noalloc call 1: R-MAT 19 ef 8 n={big.n} m={big.m} k=64 repaired", flush=True)
    assert big.n <= tgp.NOALLOC_MAX_N and big.m <= tgp.NOALLOC_MAX_M
    a, b = both(big, 64, init=tgp.INIT_BLOCK, repair=True)
    assert_same(a, b, "R-MAT 19 k=64")
    # second call, smaller, the reference's semantics
    small = tgp.gen_er(5000, 10.0, 99)
    print(f"noalloc call 2: ER n={small.n} m={small.m} k=8 reference semantics", flush=True)
    a, b = both(small, 8, init=tgp.INIT_BLOCK)
    assert_same(a, b, "ER 5000 k=8")
    # and from device-resident input
    dg = tgp.DeviceGraph(small)
    try:
        a, b = both(small, 5, init=tgp.INIT_BLOCK, repair=True, device_graph=dg.pointers())
    finally:
        dg.free()
    assert_same(a, b, "ER 5000 k=5 device input")


def test_noalloc_rejects_sizes_beyond_its_limits():
    n = tgp.NOALLOC_MAX_N + 1
    g = tgp.Graph(np.zeros(n + 1, np.int64), np.zeros(0, np.int32))
    with pytest.raises(tgp.TGPError) as ei:
        tgp.partition(g, 4, init=tgp.INIT_BLOCK, noalloc=True)
    assert ei.value.code == -1 and "tgp_partition_noalloc takes" in ei.value.detail
    small = tgp.gen_er(500, 6.0, 1)
    with pytest.raises(tgp.TGPError) as ei:
        tgp.partition(small, tgp.NOALLOC_MAX_K + 1, init=tgp.INIT_BLOCK, noalloc=True)
    assert ei.value.code == -1
