"""The synthetic configurations of gen.py (host side), checked at reduced
scale: shapes, distinct directed edges without self-loops, the request
stream's sizes, and the device generator's CPU-visible contract in bench.py."""
import numpy as np

from treesvd_amd import gen


def test_served_workload_shape_at_reduced_scale():
    w = gen.workload("served", scale_down=9)
    assert w.directed and w.transform == "LP" and w.nparts == 1024 and w.rank == 128 and w.fanout == 4
    s = np.concatenate([x[0] for x in w.snapshots]).astype(np.int64)
    d = np.concatenate([x[1] for x in w.snapshots]).astype(np.int64)
    assert (s != d).all() and s.max() < w.n and d.max() < w.n
    assert np.unique(s * w.n + d).size == s.size
    init = w.snapshots[0][0].size
    assert abs(init - 0.8 * (177_862_656 // 512)) <= 1
    sizes = np.array([x[0].size for x in w.snapshots[1:]])
    assert len(sizes) == 60 and sizes.min() >= 1
    assert set(w.subset.tolist()) <= set(np.unique(w.snapshots[0][0]).tolist())


def test_request_stream_takes_the_edges_after_the_initial_snapshot():
    w = gen.workload("served", scale_down=9)
    reqs, arr = gen.request_stream(w, 20, 10.0, seed=3, median=100, lo=10, hi=10_000)
    rest_s = np.concatenate([x[0] for x in w.snapshots[1:]])
    got = np.concatenate([r[0] for r in reqs])
    assert np.array_equal(got, rest_s[:got.size])
    assert (np.diff(arr) >= 0).all() and arr[0] == 0.0
    assert all(10 <= r[0].size <= 10_000 for r in reqs[:-1])
