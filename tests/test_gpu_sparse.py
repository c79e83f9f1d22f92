"""The n >= 1e7-node path (sparse_state pipelines) through the C ABI.

Reference: Tree_d_LP.cpp:1377-2024 with All_initialization_and_push_final
(ppr:6188-6832) and the pair-list transform (ppr:6898-7229).  Checked:
  * the batch-order insertion update (tsvd_ppr_refresh_edges) against the
    oracle's edge-by-edge restatement, from near-exact states on a tiny graph
    with vertices that get their first edge (Q8), before and after the push;
  * each snapshot on the device's own state: the stored lists hold only
    positive entries, and with one batch they are exactly the positive part
    of the pushed state; every rebuilt block equals log10(1 + (p + p^T)/rmax)
    of those lists with rows >= 30 floor(|S|/30) empty (Q11); norms and dirty
    flags follow Eqn 2 from dyn_start on; the root equals the exact FP64 merge
    tree over the device's leaves;
  * against the oracle pipeline: reserves within the push bound and the root
    singular values; one batch of source groups and several give the same
    result.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from treesvd_amd import gen
from helpers import column_mass

pytestmark = pytest.mark.gpu
ALPHA = 0.2


def q8_graph(seed):
    """40 vertices; 27-39 have no out-edge at first.  The batch gives 27-35
    their first out-edges and adds edges into them both before and after
    those (in a shuffled order), plus ordinary edges."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 40
    live = np.arange(0, 27)
    u = rng.choice(live, 160).astype(np.int32)
    v = rng.integers(0, n, 160).astype(np.int32)
    first = np.arange(27, 36)
    batch = [(f, int(rng.integers(0, n))) for f in first]              # first out-edges
    batch += [(int(rng.choice(live)), f) for f in first for _ in range(2)]   # edges into them
    batch += [(int(rng.choice(live)), int(rng.integers(0, n))) for _ in range(20)]
    batch = [batch[i] for i in rng.permutation(len(batch))]
    bu = np.array([a for a, b in batch], np.int32)
    bv = np.array([b for a, b in batch], np.int32)
    keep, bkeep = u != v, bu != bv
    return n, u[keep], v[keep], bu[bkeep], bv[bkeep]


def test_refresh_edges_matches_oracle_in_batch_order(capi):
    n, u, v, bu, bv = q8_graph(2)
    src = np.array([0, 3, 5, 8, 11, 30], np.int32)
    rmax = 1e-13
    g = capi.Graph(n, True); g.append(u, v)
    ppr = capi.PPR(g, src, ALPHA, rmax, precision=capi.TSVD_FP64)
    og = orc.Graph(n, True); og.append(u, v)
    opp = orc.PPR(n, src, True, ALPHA, rmax, fp32=False)
    alt = orc.PPR(n, src, True, ALPHA, rmax, fp32=False)
    opp.init_static(); opp.push(og)
    alt.init_static(); alt.push(og)
    for d in (0, 1):
        ppr.refresh(capi.TSVD_REFRESH_INIT, d); ppr.push(d)
        assert np.abs(ppr.p(d) - opp.p(d)).max() < 1e-11
    g.append(bu, bv); og.append(bu, bv)
    deg, former = g.degrees(0)
    assert ((former[bu] == 0) & (deg[bu] > 0)).sum() >= 5
    opp.refresh_edges(og, bu, bv)
    alt.refresh_dynamic(og)                       # the small-graph path: no Q8
    for d in (0, 1):
        ppr.refresh_edges(bu, bv, d)
        assert np.abs(ppr.p(d) - opp.p(d)).max() < 1e-10, d
        assert np.abs(ppr.r(d) - opp.r(d)).max() < 1e-10, d
    assert np.abs(opp.r(0) - alt.r(0)).max() > 1e-3   # Q8 is exercised
    opp.push(og)
    for d in (0, 1):
        ppr.push(d)
        assert np.abs(ppr.p(d) - opp.p(d)).max() < 1e-10, d
        assert np.abs(ppr.r(d) - opp.r(d)).max() < 1e-10, d


def workload(seed=3, nsrc=100):
    s, d = gen.chung_lu_edges(3000, 24000, seed, 2.1, 2.3)
    snaps = gen.shuffle_split(s, d, seed, 0.7, [2400] * 3)
    sub = gen.sample_subset(snaps[0][0], nsrc, seed)
    return 3000, snaps, sub


def lists_of(pipe, nsrc, n):
    out = []
    for d in (0, 1):
        P, R = np.zeros((nsrc, n)), np.zeros((nsrc, n))
        for i in range(nsrc):
            P[i], R[i] = pipe.source_state(i, d)
        out.append((P, R))
    return out


@pytest.mark.parametrize("prec,groups", [(1, 0), (0, 0), (0, 1)], ids=["fp64-one-batch", "fp32-one-batch",
                                                                      "fp32-two-batches"])
def test_sparse_pipeline_stagewise(capi, prec, groups):
    n, snaps, sub = workload()
    rmax, nparts, fanout, k, q, delta = 1e-6, 8, 2, 10, 6, 0.3
    pipe = capi.Pipeline(n, True, sub, ALPHA, rmax, nparts, fanout, k, q, precision=prec,
                         transform=capi.TSVD_TRANSFORM_LP, delta=delta, dyn_start=2, lazy_start=2, seed=4,
                         sparse_state=True, sparse_groups=groups)
    nsrc = sub.size
    keep = 30 * (nsrc // 30)
    starts, widths = orc.block_map(n, nparts)
    fp32 = prec == 0
    A_prev = [np.zeros((nsrc, w)) for w in widths]
    resid_prev = np.zeros(nparts)
    for it, (s, d) in enumerate(snaps, start=1):
        pipe.snapshot(s, d)
        L = lists_of(pipe, nsrc, n)
        for dd, (P, R) in enumerate(L):
            assert (P >= 0).all() and (R >= 0).all()
            deg = np.diff(pipe.graph.csr(dd)[0])
            assert (R[:, deg > 0] <= rmax * (1 + 1e-6)).all()
            if groups == 0:                      # one batch: the lists are the positive part of the state
                assert np.array_equal(P, np.maximum(pipe.ppr.p(dd), 0.0))
                assert np.array_equal(R, np.maximum(pipe.ppr.r(dd), 0.0))
        X = L[0][0] + L[1][0]                    # float p + float p^T, summed in double
        A_full = np.where(X > 0, np.log10(1.0 + X / rmax), 0.0)
        A_full[keep:] = 0.0
        a, e, r, dirty = pipe.block_norms()
        for j in range(nparts):
            Aj = A_full[:, starts[j]:starts[j] + widths[j]]
            assert a[j] == pytest.approx(np.linalg.norm(Aj), rel=1e-5 if fp32 else 1e-10)
            assert e[j] == pytest.approx(np.linalg.norm(Aj - A_prev[j]), rel=1e-4 if fp32 else 1e-9, abs=1e-6)
            want = True if it < 2 else not (resid_prev[j] + e[j] < delta * np.sqrt(2) * a[j])
            assert bool(dirty[j]) == (want and a[j] > 0), (it, j)
        for j in np.flatnonzero(dirty):
            Aj = pipe.block_matrix(j, widths[j])
            assert np.allclose(Aj, A_full[:, starts[j]:starts[j] + widths[j]], rtol=1e-6 if fp32 else 1e-13,
                               atol=1e-6 if fp32 else 1e-13)
            A_prev[j] = Aj
        resid_prev = r.copy()
        # root = exact FP64 merge tree over the device leaves
        leaves = [pipe.node(0, j) for j in range(nparts)]
        levels = orc.tree_levels(nparts, fanout)
        cur = leaves
        for lvl in range(1, len(levels)):
            nxt = []
            for p_ in range(levels[lvl]):
                ch = cur[p_ * fanout:min((p_ + 1) * fanout, len(cur))]
                if lvl == len(levels) - 1:
                    E_ref = orc.root_embedding(ch, k, "LP")
                else:
                    nxt.append(orc.merge(ch, k)[0])
            cur = nxt
        E = pipe.embedding()
        G, G_ref = E @ E.T, E_ref @ E_ref.T
        err = np.abs(G - G_ref).max() / np.abs(G_ref).max()
        assert err <= (2e-3 if fp32 else 1e-7), f"root Gram relative error {err:.3e} at snapshot {it}"
        assert not E[keep:].any()


def test_sparse_pipeline_vs_oracle_and_batching(capi):
    """Device (one batch and two) vs the oracle pipeline on the same stream.
    Reserves agree within the push bound |p_dev - p_orc| <= K rmax c_t (K
    grows with the snapshots: each stores a slightly different set of dropped
    negative residues).  Root singular values agree to the push error as the
    LP transform log10(1 + x/rmax) sees it: entries of order rmax move by
    O(1) in the log, so the bound is looser than for the reserves.  Before any
    insertion the sparse path is the dense one on the rows it keeps (sources
    do not interact), so there it must agree with a dense device pipeline
    over the first 30 floor(|S|/30) sources to the FP32 level of the dense
    path's blocks and sketch: the top half of the root sigma to 1e-3, all of
    it to 1e-2 (the smallest of the top 10 moves with the FP32 push order;
    1e-3 to 3.2e-3 seen)."""
    n, snaps, sub = workload(seed=5)
    rmax, nparts, fanout, k, q = 1e-7, 8, 2, 10, 6
    kw = dict(transform=capi.TSVD_TRANSFORM_LP, delta=0.3, dyn_start=2, lazy_start=2, seed=9)
    one = capi.Pipeline(n, True, sub, ALPHA, rmax, nparts, fanout, k, q, precision=1, sparse_state=True,
                        sparse_groups=0, **kw)
    two = capi.Pipeline(n, True, sub, ALPHA, rmax, nparts, fanout, k, q, precision=1, sparse_state=True,
                        sparse_groups=1, **kw)
    keep = 30 * (sub.size // 30)
    dense = capi.Pipeline(n, True, sub[:keep], ALPHA, rmax, nparts, fanout, k, q, precision=0, **kw)
    opipe = orc.Pipeline(n, True, sub, ALPHA, rmax, nparts, fanout, k, q, mode="LP", delta=0.3, dyn_start=2,
                         lazy_start=2, fp32=True, seed=9, nthreads=4, sparse_state=True)
    rows = []
    for it, (s, d) in enumerate(snaps, start=1):
        one.snapshot(s, d)
        two.snapshot(s, d)
        opipe.snapshot(s, d)
        if it == 1:
            dense.snapshot(s, d)
            Ed, Es = dense.embedding(), one.embedding()[:keep]
            Gd, Gs = Ed @ Ed.T, Es @ Es.T
            static_err = np.abs(Gd - Gs).max() / np.abs(Gd).max()
            rel = np.abs(dense.root_sigma() - one.root_sigma()) / dense.root_sigma()
            static_sig, static_top = rel.max(), rel[:k // 2].max()
        worst, batch_gap = 0.0, 0.0
        for dd in (0, 1):
            rp, col = one.graph.csr(dd)
            c = column_mass(n, rp, col, ALPHA)
            P_orc = opipe.ppr.p(dd)
            for i in range(0, sub.size, 7):
                p1, _ = one.source_state(i, dd)
                p2, _ = two.source_state(i, dd)
                worst = max(worst, (np.abs(p1 - P_orc[i]) / (rmax * c + 1e-12)).max())
                batch_gap = max(batch_gap, (np.abs(p1 - p2) / (rmax * c + 1e-12)).max())
        s1, s2, so = one.root_sigma(), two.root_sigma(), opipe.root_sigma
        rows.append((it, worst, batch_gap, np.abs(s1 - s2).max() / s1.max(), np.abs(s1 - so) / so))
        print(f"snapshot {it}: |p_dev - p_orc|/(rmax c) {worst:.2f}, one vs two batches {batch_gap:.2f}, "
              f"sigma one vs two {rows[-1][3]:.2e}, sigma vs oracle {np.round(rows[-1][4], 4)}")
    print(f"static snapshot, sparse vs dense device: root sigma {static_sig:.2e} (top half {static_top:.2e}), "
          f"root Gram {static_err:.2e}")
    # the singular vectors of a rank-k cut move with the sigma_k / sigma_k+1 gap
    assert static_top < 1e-3 and static_sig < 1e-2 and static_err < 2e-2
    for it, worst, batch_gap, s12, diff in rows:
        assert worst <= 4.0 * it and batch_gap <= 4.0 * it, it
        assert s12 < 1e-3, it
        assert diff[:k // 2].max() < 1e-2 and diff.max() < 3e-2, it


def test_sparse_host_batches_above_the_staging_limit(capi):
    """Host batches larger than the graph's staging limit free their device
    copy after use; the sparse-state insertion update reads that copy, so it
    must outlive the append.  With the limit at 100 edges every batch of this
    stream is above it: the run must match one with the default limit within
    the bound the batching test uses (two runs differ only in the order of
    their atomic adds, and so in which negative residues they drop)."""
    n, snaps, sub = workload(seed=7, nsrc=64)
    rmax = 1e-6
    kw = dict(transform=capi.TSVD_TRANSFORM_LP, delta=0.3, dyn_start=2, lazy_start=2, seed=5, precision=1,
              sparse_state=True)
    low = capi.Pipeline(n, True, sub, ALPHA, rmax, 8, 2, 10, 6, **kw)
    ref = capi.Pipeline(n, True, sub, ALPHA, rmax, 8, 2, 10, 6, **kw)
    low.graph.set_staging_limit(100)
    assert all(s.size > 100 for s, _ in snaps)
    for it, (s, d) in enumerate(snaps, start=1):
        low.snapshot(s, d)
        ref.snapshot(s, d)
        for dd in (0, 1):
            rp, col = ref.graph.csr(dd)
            c = column_mass(n, rp, col, ALPHA)
            for i in (0, 17, 63):
                a, b = low.source_state(i, dd)[0], ref.source_state(i, dd)[0]
                # the lists store FP32 values; each snapshot drops a slightly
                # different set of negative residues (as in the batching test)
                assert (np.abs(a - b) <= 4.0 * it * rmax * c + 4e-7 * np.abs(b) + 1e-12).all(), (it, dd, i)
        s1, s2 = low.root_sigma(), ref.root_sigma()
        assert np.abs(s1 - s2).max() / s2.max() < 1e-3
