"""The device R-MAT generator (tsvd_gen_rmat) used for the 1.47e9-edge
configuration: exactly m distinct edges, no self-loops, ids in [0, n), the
same edges for the same seed, and the quadrant masses of its top level
(a + b of the edges in the upper half of the source range before the modulo,
which at n = 2^scale is the source id's top bit)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gen_rmat_distinct_in_range_and_seeded(capi):
    scale, m = 16, 200_000
    n = 1 << scale
    s1, d1 = capi.gen_rmat(scale, n, m, 0.57, 0.19, 0.19, seed=4)
    u, v = s1.numpy(), d1.numpy()
    assert u.size == m and v.size == m
    assert (u >= 0).all() and (u < n).all() and (v >= 0).all() and (v < n).all()
    assert (u != v).all()
    keys = u.astype(np.int64) * n + v
    assert np.unique(keys).size == m
    s2, d2 = capi.gen_rmat(scale, n, m, 0.57, 0.19, 0.19, seed=4)
    assert np.array_equal(s2.numpy(), u) and np.array_equal(d2.numpy(), v)
    s3, _ = capi.gen_rmat(scale, n, m, 0.57, 0.19, 0.19, seed=5)
    assert not np.array_equal(s3.numpy(), u)
    # top level: source in the upper half with probability c + d = 0.24,
    # target in the right half with b + d = 0.24 (distinct-edge dedup removes
    # mostly heavy top-left pairs, so allow a few percent)
    top_u = (u >= n // 2).mean()
    top_v = (v >= n // 2).mean()
    assert 0.22 < top_u < 0.30 and 0.22 < top_v < 0.30, (top_u, top_v)
    # ids modulo n fold the range
    s4, d4 = capi.gen_rmat(scale, 40_000, 50_000, 0.57, 0.19, 0.19, seed=1)
    u4, v4 = s4.numpy(), d4.numpy()
    assert u4.max() < 40_000 and v4.max() < 40_000 and (u4 != v4).all()
    assert np.unique(u4.astype(np.int64) * 40_000 + v4).size == 50_000
    for x in (s1, d1, s2, d2, s3, s4, d4):
        x.free()


def test_gen_rmat_rejects_bad_arguments(capi):
    with pytest.raises(capi.TreeSVDError):
        capi.gen_rmat(10, 2000, 100, 0.57, 0.19, 0.19)          # n > 2^scale
    with pytest.raises(capi.TreeSVDError):
        capi.gen_rmat(10, 1024, 100, 0.7, 0.2, 0.2)             # a + b + c >= 1
    with pytest.raises(capi.TreeSVDError):
        capi.gen_rmat(4, 16, 10_000, 0.57, 0.19, 0.19)          # more edges than distinct pairs
