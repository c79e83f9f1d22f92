"""The plain per-snapshot method (tsvd_pipeline_*) through the C ABI.

Each stage is checked on the device's own inputs, so push-order differences
(bounded in test_gpu_push.py) do not enter the comparison:
  * block matrices = the reference transform of the device reserves
    (ppr:2527-2638), norms and dirty flags = Eqn 2 on those matrices;
  * every dirty leaf = the oracle's FP64 frPCA (frpca.c:239-326) of the same
    block with the same sketch (compared as U S^2 U^T, rotation-free);
  * merges and root = the oracle's exact FP64 SVD tree (Tree_u_NC.cpp:723-850)
    over the device leaves;
and the whole path is compared with the FP64 oracle pipeline and with the
dense FP64 SVD of the full matrix."""
import numpy as np
import pytest

from oracle import oracle as orc
from treesvd_amd import gen

pytestmark = pytest.mark.gpu


def small_workload(directed, seed=3):
    if directed:
        s, d = gen.chung_lu_edges(3000, 24000, seed, 2.1, 2.3)
        n = 3000
    else:
        s, d, n = gen.rmat_edges(12, 24000, seed)
    snaps = gen.shuffle_split(s, d, seed, 0.7, [2400] * 3)
    sub = gen.sample_subset(snaps[0][0], 80, seed)
    return n, snaps, sub


def transform_of(pipe, n, directed, rmax, mode):
    p = pipe.ppr.p(0)
    if directed:
        p = p + pipe.ppr.p(1)
    return orc.transform(p, rmax, mode)


def active_columns(pipe, n, directed, sources):
    deg0, _ = pipe.graph.degrees(0)
    deg = deg0 + (pipe.graph.degrees(1)[0] if directed else 0)
    act = deg > 0
    act[sources] = True
    return act


@pytest.mark.parametrize("directed,prec", [(False, 0), (False, 1), (True, 0)])
def test_block_norms_of_changed_items_equal_a_full_recompute(capi, directed, prec):
    """The norm pass re-reads only the (vertex, source group) items whose
    reserves the insertion update or the push changed, and keeps the others'
    sums.  With insertions of a few dozen edges most items stay untouched, so
    ||A_new|| and ||A_new - A_old|| (ppr:2649-2732) are checked against the
    transform of the whole device state at every snapshot, across lazy-test
    rebuilds (A_old replaced only for dirty blocks)."""
    if directed:
        s, d = gen.chung_lu_edges(8000, 60000, 5, 2.1, 2.3)
        n = 8000
    else:
        s, d, n = gen.rmat_edges(14, 90000, 5)
    snaps = gen.shuffle_split(s, d, 5, 0.8, [60] * 6 + [3000])
    sub = gen.sample_subset(snaps[0][0], 150, 5)          # 3 source groups, the last ragged
    rmax, nparts, delta = 1e-5, 16, 0.03
    pipe = capi.Pipeline(n, directed, sub, 0.2, rmax, nparts, 2, 8, 4, precision=prec, delta=delta, seed=9)
    starts, widths = orc.block_map(n, nparts)
    fp32 = prec == 0
    A_prev = [np.zeros((sub.size, w)) for w in widths]
    items = n * 3
    partial = 0
    for it, (a_s, a_d) in enumerate(snaps, start=1):
        st = pipe.snapshot(a_s, a_d)
        A_full = transform_of(pipe, n, directed, rmax, "NC")
        a, e, _, dirty = pipe.block_norms()
        for j in range(nparts):
            Aj = A_full[:, starts[j]:starts[j] + widths[j]]
            assert a[j] == pytest.approx(np.linalg.norm(Aj), rel=1e-5 if fp32 else 1e-10), (it, j)
            assert e[j] == pytest.approx(np.linalg.norm(Aj - A_prev[j]), rel=1e-4 if fp32 else 1e-9,
                                         abs=1e-6), (it, j)
        for j in np.flatnonzero(dirty):
            A_prev[j] = A_full[:, starts[j]:starts[j] + widths[j]].copy()
        assert 0 < st["norm_items"] <= items
        partial += st["norm_items"] < items
        print(f"snapshot {it}: {st['norm_items']} of {items} items re-read, {int(dirty.sum())} dirty", flush=True)
    assert st["snapshot"] == len(snaps)
    assert partial >= 4, "small insertions should leave most items untouched"
    pipe.close()


@pytest.fixture(params=["density-rule", "all-sparse", "all-dense"])
def leaf_products(request, monkeypatch):
    """Leaf products by the library's density rule, from sparse CSR/CSC
    copies for every dirty block, or by dense GEMM for every dirty block
    (TSVD_SPARSE_DENSITY, read when the pipeline is created)."""
    value = {"density-rule": None, "all-sparse": "1.01", "all-dense": "0"}[request.param]
    if value is None:
        monkeypatch.delenv("TSVD_SPARSE_DENSITY", raising=False)
    else:
        monkeypatch.setenv("TSVD_SPARSE_DENSITY", value)
    return request.param


@pytest.mark.parametrize("directed,prec,mode", [(False, 1, "NC"), (False, 0, "NC"), (True, 0, "NC"),
                                                 (False, 0, "LP")])
def test_pipeline_stagewise(capi, directed, prec, mode, leaf_products):
    n, snaps, sub = small_workload(directed)
    rmax, nparts, fanout, k, q, delta = 1e-5, 8, 2, 12, 6, 0.12
    seed = 42
    pipe = capi.Pipeline(n, directed, sub, 0.2, rmax, nparts, fanout, k, q, precision=prec,
                         transform=capi.TSVD_TRANSFORM_NC if mode == "NC" else capi.TSVD_TRANSFORM_LP,
                         delta=delta, seed=seed)
    starts, widths = orc.block_map(n, nparts)
    fp32 = prec == 0
    rtol = 2e-4 if fp32 else 1e-8
    A_prev = [np.zeros((sub.size, w)) for w in widths]
    resid_prev = np.zeros(nparts)
    for it, (s, d) in enumerate(snaps, start=1):
        st = pipe.snapshot(s, d)
        assert st["snapshot"] == it
        A_full = transform_of(pipe, n, directed, rmax, mode)
        a, e, r, dirty = pipe.block_norms()
        act = active_columns(pipe, n, directed, sub)
        # norms and the lazy test (ppr:2699-2726), all blocks dirty before lazy_start = 2
        for j in range(nparts):
            Aj = A_full[:, starts[j]:starts[j] + widths[j]]
            assert a[j] == pytest.approx(np.linalg.norm(Aj), rel=1e-5 if fp32 else 1e-10)
            assert e[j] == pytest.approx(np.linalg.norm(Aj - A_prev[j]), rel=1e-4 if fp32 else 1e-9, abs=1e-6)
            want = True if it < 2 else not (resid_prev[j] + e[j] < delta * np.sqrt(2) * a[j])
            assert bool(dirty[j]) == (want and a[j] > 0)
        for j in np.flatnonzero(dirty):
            Aj = pipe.block_matrix(j, widths[j])
            assert np.allclose(Aj, A_full[:, starts[j]:starts[j] + widths[j]], rtol=1e-5 if fp32 else 1e-12,
                               atol=1e-6 if fp32 else 1e-12)
            A_prev[j] = Aj
            # leaf j vs the oracle frPCA of the same compacted block and sketch
            cols = np.flatnonzero(act[starts[j]:starts[j] + widths[j]])
            assert np.array_equal(pipe.block_columns(j), starts[j] + cols)
            Ac = Aj[:, cols]
            om = orc.sketch(orc.leaf_seed(seed, j), cols.size, k + 5, fp32=fp32)
            U, S, V = orc.frpca(Ac, k, q, omega=om)
            leaf = pipe.node(0, j)
            ref = (U * S) @ (U * S).T
            err = np.abs(leaf @ leaf.T - ref).max() / np.abs(ref).max()
            assert err <= rtol * 10, f"leaf {j} snapshot {it}: relative Gram error {err:.3e}"
            res = np.linalg.norm((U * S) @ V.T - Ac)
            assert r[j] == pytest.approx(res, rel=5e-3 if fp32 else 1e-6, abs=1e-6 * np.linalg.norm(Ac))
        resid_prev = r.copy()
        # merges + root over the device leaves, exact FP64 SVDs
        leaves = [pipe.node(0, j) for j in range(nparts)]
        levels = orc.tree_levels(nparts, fanout)
        cur = leaves
        for lvl in range(1, len(levels)):
            nxt = []
            for p_ in range(levels[lvl]):
                ch = cur[p_ * fanout:min((p_ + 1) * fanout, len(cur))]
                if lvl == len(levels) - 1:
                    E_ref = orc.root_embedding(ch, k, mode)
                    _, _, S_ref = orc.merge(ch, k)
                else:
                    nxt.append(orc.merge(ch, k)[0])
            cur = nxt
        E = pipe.embedding()
        G, G_ref = E @ E.T, E_ref @ E_ref.T
        err = np.abs(G - G_ref).max() / np.abs(G_ref).max()
        assert err <= rtol * 10, f"root Gram relative error {err:.3e} at snapshot {it}"
        serr = np.abs(pipe.root_sigma() - S_ref).max() / S_ref.max()
        assert serr <= rtol * 10, f"root sigma relative error {serr:.3e} at snapshot {it}"


def test_pipeline_vs_oracle_pipeline_and_dense_svd(capi):
    """Whole path, FP64: device vs the oracle pipeline (its own FIFO push,
    same sketches) and both vs the dense FP64 SVD of the full matrix."""
    n, snaps, sub = small_workload(False, seed=5)
    rmax, nparts, fanout, k, q = 1e-7, 8, 2, 12, 6
    pipe = capi.Pipeline(n, False, sub, 0.2, rmax, nparts, fanout, k, q, precision=1, seed=9)
    opipe = orc.Pipeline(n, False, sub, 0.2, rmax, nparts, fanout, k, q, fp32=False, seed=9, nthreads=4)
    for s, d in snaps:
        pipe.snapshot(s, d)
        opipe.snapshot(s, d)
    A = transform_of(pipe, n, False, rmax, "NC")
    s_dense = np.linalg.svd(A, compute_uv=False)[:k]
    s_gpu = pipe.root_sigma()
    s_orc = opipe.root_sigma
    err_gpu = np.abs(s_gpu - s_dense) / s_dense
    err_orc = np.abs(s_orc - s_dense) / s_dense
    diff = np.abs(s_gpu - s_orc) / s_orc
    msg = f"vs dense: device {np.round(err_gpu, 4)}, oracle {np.round(err_orc, 4)}; device vs oracle {np.round(diff, 5)}"
    print(msg)
    # the device error tracks the reference algorithm's own error, and the two
    # agree to the push error (different push order, same sketches)
    assert err_gpu.max() <= err_orc.max() + 1e-3, msg
    assert diff[:k // 2].max() < 1e-3 and diff.max() < 2e-2, msg


@pytest.mark.parametrize("directed", [False, True])
def test_dense_gram_gram_apply_and_iterative_oracle(capi, directed):
    """The accuracy oracles on the device: the FP64 Gram A A^T and the Gram-free
    product A (A^T Y) both equal the host products of the transformed device
    reserves; subspace iteration through gram_apply reaches numpy's eigh of
    that Gram (top-k eigenvalues, the spanned subspace and its certificate)."""
    import bench
    n, snaps, sub = small_workload(directed, seed=6)
    rmax, k = 1e-6, 10
    pipe = capi.Pipeline(n, directed, sub, 0.2, rmax, 4, 2, k, 6, precision=1, seed=3)
    for s, d in snaps[:2]:
        pipe.snapshot(s, d)
    A = transform_of(pipe, n, directed, rmax, "NC")
    G = A @ A.T
    scale = np.abs(G).max()
    assert np.abs(pipe.dense_gram() - G).max() <= 1e-12 * scale * n ** 0.5
    Y = np.random.Generator(np.random.PCG64(0)).standard_normal((sub.size, 7))
    assert np.abs(pipe.gram_apply(Y) - G @ Y).max() <= 1e-12 * scale * n ** 0.5 * np.abs(Y).max() * sub.size
    lam, V = np.linalg.eigh(G)
    lam, V = lam[::-1][:k], V[:, ::-1][:, :k]
    it = bench.DenseOracle(pipe, k, np.sqrt((A.reshape(A.shape[0], -1) ** 2).sum()).reshape(1), gram_limit=0,
                           tol=1e-10)
    assert it.bound < 1e-10
    assert np.abs(it.lam - lam).max() <= 1e-8 * lam[0]
    assert abs(it.tr - np.trace(G)) <= 1e-10 * np.trace(G)
    # same subspace: 1 - ||V^T U||^2 / k ~ 0
    assert 1.0 - np.linalg.norm(V.T @ it.U) ** 2 / k < 1e-8
    dense = bench.DenseOracle(pipe, k, None)
    assert dense.G is not None and np.abs(dense.lam - lam).max() <= 1e-10 * lam[0]


@pytest.mark.parametrize("directed,mode", [(False, "NC"), (True, "LP")])
def test_block_transform_entry_point(capi, directed, mode):
    """tsvd_pipeline_block_transform, the per-block-range replacement of
    Log_sparse_matrix_entries_with_norm_computation[_LP] (ppr:2649-2826) and,
    with test = 0, of Log_sparse_matrix_entries[_LP] (ppr:2527-2638).  After
    two snapshots, a third insertion goes through the graph and push handles
    only (append, dynamic refresh, push), as a driver calling the pieces
    would; the block transform over a sub-range then gives the norms of the
    oracle's transform of the device reserves, the Eqn-2 record (-1 clean,
    iter dirty) on those norms and the stored residual, replaces exactly the
    dirty blocks' stored matrices and leaves blocks outside the range alone."""
    n, snaps, sub = small_workload(directed)
    rmax, nparts, k, delta = 1e-5, 8, 12, 0.12
    pipe = capi.Pipeline(n, directed, sub, 0.2, rmax, nparts, 2, k, 6, precision=1, delta=delta, seed=3,
                         transform=capi.TSVD_TRANSFORM_NC if mode == "NC" else capi.TSVD_TRANSFORM_LP)
    for s, d in snaps[:2]:
        pipe.snapshot(s, d)
    pipe.graph.append(*snaps[2])
    for d in ((0, 1) if directed else (0,)):
        pipe.ppr.refresh(capi.TSVD_REFRESH_DYNAMIC, d)
        pipe.ppr.push(d)
    A_full = transform_of(pipe, n, directed, rmax, mode)
    starts, widths = orc.block_map(n, nparts)
    resid = pipe.block_norms()[2]
    old = [pipe.block_matrix(j, widths[j]) for j in range(nparts)]
    j0, j1, it = 2, 7, 3
    a, e, rec = pipe.block_transform(j0, j1, it, delta, test=True)
    assert a.size == e.size == rec.size == j1 - j0
    for i, j in enumerate(range(j0, j1)):
        Aj = A_full[:, starts[j]:starts[j] + widths[j]]
        _, an, en = orc.dirty_test(Aj, old[j], resid[j], delta)
        assert a[i] == pytest.approx(an, rel=1e-10) and e[i] == pytest.approx(en, rel=1e-9, abs=1e-12)
        dirty = not (resid[j] + e[i] < delta * np.sqrt(2.0) * a[i])
        assert rec[i] == (it if dirty else -1)
        now = pipe.block_matrix(j, widths[j])
        if dirty:
            assert np.allclose(now, Aj, rtol=1e-12, atol=1e-12)
        else:
            assert np.array_equal(now, old[j])
    for j in list(range(j0)) + list(range(j1, nparts)):
        assert np.array_equal(pipe.block_matrix(j, widths[j]), old[j])
    # test = 0: every block recorded and replaced, norms against what is stored now
    stored = [pipe.block_matrix(j, widths[j]) for j in range(nparts)]
    a0, e0, rec0 = pipe.block_transform(0, nparts, 4, delta, test=False)
    assert (rec0 == 4).all()
    for j in range(nparts):
        Aj = A_full[:, starts[j]:starts[j] + widths[j]]
        assert e0[j] == pytest.approx(np.linalg.norm(Aj - stored[j]), rel=1e-9, abs=1e-12)
        assert np.allclose(pipe.block_matrix(j, widths[j]), Aj, rtol=1e-12, atol=1e-12)
    with pytest.raises(capi.TreeSVDError):
        pipe.block_transform(3, nparts + 1, 5, delta)
