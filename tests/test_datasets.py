"""Dataset tools (treesvd_amd.datasets) against the rules of the reference's
programs, checked by brute force on small inputs and by round trips through
the drivers' readers in the C ABI (host code only, no GPU).

Rules pinned here (Random_Sample_Points_LP.cpp:80-160; gen_subset_LP_data_u
.cpp:61-146 with Graph_dynamic.h:246-309, 665-723; split_edges_NC.cpp:57-118):
the subset is distinct, drawn among the first snapshot's edge sources and
written sorted; the hold-out moves int(percent * m) edges to the front by a
partial Fisher-Yates shuffle, keeps those with a subset end as positives with
the subset end first (the from end when both are), writes the rest as the
training snapshot, and draws as many negatives, a subset vertex and a vertex
not adjacent to it in the whole graph (out-adjacency when directed); batches
are consecutive runs of exactly `group` edges, at most `max_batches`, the
remainder dropped.  The reference draws with rand(), so the draws themselves
are parity unpinned; the rules are not."""
import numpy as np
import pytest

from treesvd_amd import datasets as ds


class FixedRng:
    """integers(k) returns the next scripted value modulo k."""

    def __init__(self, values):
        self.values = list(values)

    def integers(self, k):
        return self.values.pop(0) % k


def test_partial_shuffle_is_the_references_swap_loop():
    src = np.arange(10, dtype=np.int32)
    dst = src + 100
    script = [3, 0, 7, 2]
    s, d, m2 = ds.partial_shuffle(src, dst, 0.45, FixedRng(script))
    assert m2 == 4
    want = list(range(10))
    for i, r in enumerate(script):                 # for (i < m2) swap(i, i + rand() % (m - i))
        j = i + r % (10 - i)
        want[i], want[j] = want[j], want[i]
    assert s.tolist() == want and (d - s == 100).all()
    assert sorted(s.tolist()) == list(range(10))


def test_sample_subset_distinct_sorted_from_first_sources():
    rng = np.random.Generator(np.random.PCG64(1))
    first_src = rng.integers(0, 500, 3000).astype(np.int32)
    sub = ds.sample_subset(first_src, 64, seed=3)
    assert sub.size == 64 and np.unique(sub).size == 64
    assert (np.diff(sub) > 0).all()
    assert np.isin(sub, first_src).all()
    assert np.array_equal(sub, ds.sample_subset(first_src, 64, seed=3))
    with pytest.raises(ValueError):
        ds.sample_subset(np.array([1, 1, 2], np.int32), 3)


@pytest.mark.parametrize("directed", [False, True])
def test_lp_holdout_rules(directed):
    rng = np.random.Generator(np.random.PCG64(7))
    n = 60
    snaps = []
    for m in (200, 80, 120):
        s = rng.integers(0, n, m).astype(np.int32)
        d = rng.integers(0, n, m).astype(np.int32)
        keep = s != d
        snaps.append((s[keep], d[keep]))
    subset = np.array([2, 5, 11, 17, 23, 40], np.int32)
    percent = 0.3
    train, pos, neg = ds.lp_holdout(snaps, subset, n, percent, directed=directed, seed=5)
    # replay the same draws to find each snapshot's test part
    rr = np.random.Generator(np.random.PCG64(5))
    in_s = np.isin(np.arange(n), subset)
    want_pos = []
    for (s, d), (ts, td) in zip(snaps, train):
        ps, pd, m2 = ds.partial_shuffle(s, d, percent, rr)
        assert m2 == int(percent * s.size)
        assert np.array_equal(ts, ps[m2:]) and np.array_equal(td, pd[m2:])
        for a, b in zip(ps[:m2], pd[:m2]):
            if in_s[a]:
                want_pos.append((a, b))
            elif in_s[b]:
                want_pos.append((b, a))
        # training + test candidates = the snapshot, as a multiset of pairs
        assert sorted(zip(s.tolist(), d.tolist())) == sorted(zip(ps.tolist(), pd.tolist()))
    assert [tuple(x) for x in pos.tolist()] == [(int(a), int(b)) for a, b in want_pos]
    assert neg.shape == pos.shape and pos.shape[0] > 0
    adj = set()
    for s, d in snaps:
        for a, b in zip(s.tolist(), d.tolist()):
            adj.add((a, b))
            if not directed:
                adj.add((b, a))
    for i, j in neg.tolist():
        assert in_s[i] and 0 <= j < n and (i, j) not in adj


def test_split_batches_exact_runs_remainder_dropped():
    snaps = [(np.arange(25, dtype=np.int32), np.arange(25, dtype=np.int32) + 1000),
             (np.arange(25, 47, dtype=np.int32), np.arange(25, 47, dtype=np.int32) + 1000)]
    b = ds.split_batches(snaps, group=10, max_batches=100)
    assert [x.size for x, _ in b] == [10, 10, 10, 10]
    assert np.array_equal(np.concatenate([x for x, _ in b]), np.arange(40))
    assert len(ds.split_batches(snaps, group=10, max_batches=2)) == 2
    assert ds.split_batches(snaps, group=100) == []


def test_written_files_read_back_through_the_drivers_readers(capi, tmp_path):
    """Files as the reference's tools write them read back with the drivers'
    readers: the edge reader's duplicate last edge (Q1) with dup_last, the
    subset file as integers, config.txt as getline lines."""
    src = np.array([4, 0, 9], np.int32)
    dst = np.array([1, 7, 2], np.int32)
    ds.write_edges(str(tmp_path / "s1.txt"), src, dst)
    s, d = capi.read_edge_file(str(tmp_path / "s1.txt"), dup_last=True)
    assert s.tolist() == [4, 0, 9, 9] and d.tolist() == [1, 7, 2, 2]
    s, d = capi.read_edge_file(str(tmp_path / "s1.txt"), dup_last=False)
    assert s.tolist() == [4, 0, 9] and d.tolist() == [1, 7, 2]
    ds.write_subset(str(tmp_path / "sub.txt"), np.array([3, 8, 20], np.int32))
    assert capi.read_subset_file(str(tmp_path / "sub.txt")).tolist() == [3, 8, 20]
    ds.write_config(str(tmp_path / "config.txt"), "sub.txt", ["s1.txt", "s2.txt"])
    assert capi.read_lines(str(tmp_path / "config.txt")) == ["sub.txt", "s1.txt", "s2.txt"]
    ds.write_pairs(str(tmp_path / "pos.txt"), np.array([[3, 1], [8, 2]], np.int32))
    s, d = capi.read_edge_file(str(tmp_path / "pos.txt"), dup_last=False)
    assert s.tolist() == [3, 8] and d.tolist() == [1, 2]
