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


@pytest.mark.parametrize("directed,prec,mode", [(False, 1, "NC"), (False, 0, "NC"), (True, 0, "NC"),
                                                 (False, 0, "LP")])
def test_pipeline_stagewise(capi, directed, prec, mode):
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
