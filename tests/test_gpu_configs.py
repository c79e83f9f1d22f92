"""Parity at the shapes of the measured configurations, through the C ABI.

The configurations (131k-node plain graph; 2.1M-node batched R-MAT;
2.7M-node directed Chung-Lu with 512 blocks; the 6.2M-node served graph;
the 41.6M-node graph) take code paths the small tests do not: rank 128
(the wide MFMA tile, l = 133), fan-out 4, the ragged 512-leaf tree (512 is
not a power of 4, so level 4 has 2 nodes and the root 2 children), |S| in
the thousands.  What is checked:

* the 131,072-node plain configuration in full against the oracle pipeline
  (Tree_u_NC.cpp:507-866 restated): reserves of all 256 sources within the
  push bound, the same dirty blocks, every dirty leaf equal to the oracle's
  frPCA of the device's block with the same sketch (frpca.c:239-326), and
  the root singular values of the two pipelines;
* the batched and precision configurations on their own graphs (initial
  snapshot + 2 insertions): 32 sampled sources against the oracle's FIFO
  push (ppr:1394-1498, 2035-2121, 2945-3030, 3356-3437, 3811-3898,
  4133-4228), termination and mass conservation, sampled dirty leaves at
  rank 128 against the oracle's frPCA, and every merge of the tree
  (Tree_u_NC.cpp:238-262, 723-850) against the FP64 spectrum of its
  children;
* fan-out-4 ragged trees at rank 128 on small graphs, stage by stage;
* one GPU's share of the served and the 41.6M-node configurations
  (source-sharded), by properties that hold at any size: termination,
  Sum_v p + Sum_v r = 1 per source (both the push and the insertion update
  conserve it), p >= 0, and every merge of the tree.

A merge is checked against what the mathematics fixes rather than a second
SVD routine: with C = [children] and sigma the top-k square roots of
eigvalsh(C^T C) (FP64, LAPACK), the parent P = U_k Sigma_k must have
orthogonal columns with norms sigma and satisfy C C^T P = P Sigma^2.
"""
import time

import numpy as np
import pytest

from oracle import oracle as orc
from treesvd_amd import gen
from helpers import column_mass

pytestmark = pytest.mark.gpu
ALPHA = 0.2


def say(msg):
    print(msg, flush=True)


def pipeline(capi, w, seed, **kw):
    return capi.Pipeline(w.n, w.directed, w.subset, w.alpha, w.rmax, w.nparts, w.fanout, w.rank, w.q,
                         precision=capi.TSVD_FP64 if w.fp64_state else capi.TSVD_FP32,
                         transform=capi.TSVD_TRANSFORM_NC if w.transform == "NC" else capi.TSVD_TRANSFORM_LP,
                         delta=w.delta, dyn_start=2, lazy_start=2, slack=0.1, seed=seed,
                         merge_bf16x3=w.merge_bf16x3, **kw)


def merge_error(P, children):
    """Largest of: column-norm error against the top-k singular values of
    C = [children], off-diagonal of P^T P, and the residual of
    C C^T P = P Sigma^2, each relative to sigma_max (squared where it is)."""
    C = np.concatenate(children, axis=1)
    k = P.shape[1]
    lam = np.linalg.eigvalsh(C.T @ C)[::-1][:k]
    sig = np.sqrt(np.clip(lam, 0.0, None))
    if k > sig.size:
        sig = np.concatenate([sig, np.zeros(k - sig.size)])
    smax = max(sig[0], 1e-300)
    nrm = np.linalg.norm(P, axis=0)
    G = P.T @ P
    off = np.abs(G - np.diag(np.diag(G))).max()
    R = C @ (C.T @ P) - P * (sig ** 2)[None, :]
    return max(np.abs(nrm - sig).max() / smax, off / smax ** 2,
               np.linalg.norm(R) / (smax ** 2 * max(np.linalg.norm(P), 1e-300)))


def tree_nodes(pipe, nparts, fanout):
    levels = orc.tree_levels(nparts, fanout)
    nodes = [[pipe.node(lvl, i) for i in range(levels[lvl])] for lvl in range(len(levels) - 1)]
    return levels, nodes


def check_tree(pipe, nparts, fanout, k, mode, tol, per_level=None, seed=0):
    """Every internal node (or, with per_level, that many random nodes of
    each level), and the root embedding, against its device children.
    Returns the largest error seen."""
    levels, nodes = tree_nodes(pipe, nparts, fanout)
    rng = np.random.Generator(np.random.PCG64(seed))
    worst = 0.0
    for lvl in range(1, len(levels) - 1):
        picks = range(levels[lvl])
        if per_level is not None and levels[lvl] > per_level:
            picks = sorted(rng.choice(levels[lvl], per_level, replace=False))
        for p_ in picks:
            ch = nodes[lvl - 1][p_ * fanout:min((p_ + 1) * fanout, levels[lvl - 1])]
            err = merge_error(nodes[lvl][p_], ch)
            assert err <= tol, f"node ({lvl}, {p_}): merge error {err:.3e}"
            worst = max(worst, err)
    top = len(levels) - 1
    ch = nodes[top - 1] if top > 0 else nodes[0]
    sig = pipe.root_sigma()
    E = pipe.embedding()
    P = E * np.sqrt(np.abs(sig))[None, :] if mode == "NC" else E * sig[None, :]
    err = merge_error(P, ch)
    assert err <= tol, f"root: merge error {err:.3e}"
    return max(worst, err)


def leaf_error(pipe, j, start, width, seed, k, q, fp32):
    """Leaf j (U S on the device) against the oracle's FP64 frPCA of the
    device's own block over its stored columns, with the same sketch,
    compared as U S^2 U^T (free of rotations within equal singular values)."""
    Aj = pipe.block_matrix(j, width)
    cols = pipe.block_columns(j) - start
    Ac = Aj[:, cols]
    del Aj
    om = orc.sketch(orc.leaf_seed(seed, j), cols.size, k + 5, fp32=fp32)
    U, S, _ = orc.frpca(Ac, k, q, omega=om)
    leaf = pipe.node(0, j)
    ref = (U * S) @ (U * S).T
    return np.abs(leaf @ leaf.T - ref).max() / np.abs(ref).max()


# --------------------------------------------------------------- tree shapes
def shape_workload(directed, n, m, nsrc, seed):
    if directed:
        s, d = gen.chung_lu_edges(n, m, seed, 2.1, 2.3)
    else:
        s, d, n = gen.rmat_edges(int(np.log2(n)), m, seed)
    snaps = gen.shuffle_split(s, d, seed, 0.7, [m // 10] * 2)
    return n, snaps, gen.sample_subset(snaps[0][0], nsrc, seed)


@pytest.mark.parametrize("directed,n,m,nparts,k,mode", [
    (False, 1 << 15, 400_000, 40, 128, "NC"),       # 40 -> 10 -> 3 -> 1, rank 128 (wide tile)
    (True, 1 << 16, 600_000, 512, 16, "NC"),        # the 512-leaf tree: 512 -> 128 -> 32 -> 8 -> 2 -> 1
    (False, 1 << 15, 400_000, 21, 96, "LP"),        # 21 -> 6 -> 2 -> 1, rank 96
], ids=["40-leaves-rank128", "512-leaves-directed", "21-leaves-rank96-LP"])
def test_fanout4_ragged_trees(capi, directed, n, m, nparts, k, mode):
    n, snaps, sub = shape_workload(directed, n, m, 256, 5)
    rmax, fanout, q, seed = 1e-6, 4, 6, 17
    pipe = capi.Pipeline(n, directed, sub, ALPHA, rmax, nparts, fanout, k, q, precision=capi.TSVD_FP32,
                         transform=capi.TSVD_TRANSFORM_NC if mode == "NC" else capi.TSVD_TRANSFORM_LP,
                         delta=0.3, seed=seed)
    starts, widths = orc.block_map(n, nparts)
    try:
        for it, (s, d) in enumerate(snaps, start=1):
            st = pipe.snapshot(s, d)
            dirty = np.flatnonzero(pipe.block_norms()[3])
            assert st["ndirty"] == dirty.size
            worst_leaf = 0.0
            for j in dirty:
                err = leaf_error(pipe, j, starts[j], widths[j], seed, k, q, True)
                assert err <= 2e-3, f"leaf {j} at snapshot {it}: {err:.3e}"
                worst_leaf = max(worst_leaf, err)
            worst = check_tree(pipe, nparts, fanout, k, mode, 2e-4)
            say(f"snapshot {it}: {dirty.size} dirty leaves (worst {worst_leaf:.2e}), tree {worst:.2e}")
    finally:
        pipe.close()


# ------------------------------------------------------------ configuration 1
def test_config_plain_in_full_against_the_oracle_pipeline(capi):
    """131,072-node R-MAT, 1M edges (70% + 3 x 10%), 256 sources,
    rmax 1e-6, 16 blocks, rank 64, q 6, fan-out 2, FP64.  The oracle
    pipeline pushes in FIFO order, the device in synchronous rounds: their
    reserves differ within the push bound, so their root singular values are
    compared through the exact spectra of their matrices (Weyl), while the
    device's root equals the reference algorithm run on the device's own
    blocks."""
    t0 = time.time()
    w = gen.workload("plain")
    seed = 9
    pipe = pipeline(capi, w, seed)
    opipe = orc.Pipeline(w.n, False, w.subset, w.alpha, w.rmax, w.nparts, w.fanout, w.rank, w.q, fp32=False,
                         seed=seed, nthreads=16, delta=w.delta)
    starts, widths = orc.block_map(w.n, w.nparts)
    try:
        for it, (s, d) in enumerate(w.snapshots, start=1):
            pipe.snapshot(s, d)
            opipe.snapshot(s, d)
            p, r = pipe.ppr.p(), pipe.ppr.r()
            op, orr = opipe.ppr.p(0), opipe.ppr.r(0)
            rp, col = pipe.graph.csr()
            deg = np.diff(rp)
            assert (r[:, deg > 0] <= w.rmax * (1 + 1e-9)).all()
            assert np.abs(p.sum(1) + r.sum(1) - 1.0).max() < 1e-9          # mass conservation
            c = column_mass(w.n, rp, col, ALPHA)
            # invariant: p_dev - p_orc = (r_orc - r_dev) Pi, so per source
            # |p_dev(t) - p_orc(t)| <= max_v |r_orc(v) - r_dev(v)| c_t
            dr = np.abs(orr - r).max(1)
            ratio = (np.abs(p - op) / (dr[:, None] * c[None, :] + 1e-15)).max()
            assert ratio <= 1.0 + 1e-6, ratio
            dirty = np.flatnonzero(pipe.block_norms()[3])
            assert sorted(dirty.tolist()) == sorted(opipe.last_dirty)
            worst = 0.0
            for j in dirty:
                err = leaf_error(pipe, j, starts[j], widths[j], seed, w.rank, w.q, False)
                assert err <= 1e-6, f"leaf {j}: {err:.3e}"
                worst = max(worst, err)
            tree = check_tree(pipe, w.nparts, w.fanout, w.rank, "NC", 1e-9)
            # the reference algorithm (the oracle's frPCA of every stored block
            # with its sketch, then exact merges) on the device's own blocks
            leaves = []
            for j in range(w.nparts):
                Aj = pipe.block_matrix(j, widths[j])
                cols = pipe.block_columns(j) - starts[j]
                if cols.size == 0 or not Aj.any():
                    leaves.append(np.zeros((w.subset.size, w.rank)))
                    continue
                om = orc.sketch(orc.leaf_seed(seed, j), cols.size, w.rank + 5)
                U, S, _ = orc.frpca(Aj[:, cols], w.rank, w.q, omega=om)
                leaves.append(U * S)
            cur = leaves
            while len(cur) > 2:
                cur = [orc.merge(cur[i:i + w.fanout], w.rank)[0] for i in range(0, len(cur), w.fanout)]
            s_ref = orc.merge(cur, w.rank)[2]
            s_dev, s_orc = pipe.root_sigma(), opipe.root_sigma
            gap = np.abs(s_dev - s_ref) / s_ref
            assert gap.max() < 1e-6, gap.max()
            # device and oracle pipelines push in different orders, so their
            # matrices differ by the push error; the exact spectra of the two
            # current matrices obey Weyl's bound |s_i(A) - s_i(B)| <= ||A - B||_2
            A_dev = orc.transform(p, w.rmax, "NC")
            A_orc = orc.transform(opipe.matrix(), w.rmax, "NC")
            e_dev = np.sqrt(np.clip(np.linalg.eigvalsh(A_dev @ A_dev.T)[::-1], 0, None))
            e_orc = np.sqrt(np.clip(np.linalg.eigvalsh(A_orc @ A_orc.T)[::-1], 0, None))
            D = A_dev - A_orc
            weyl = np.sqrt(max(np.linalg.eigvalsh(D @ D.T)[-1], 0.0))
            assert np.abs(e_dev - e_orc).max() <= weyl * (1 + 1e-9) + 1e-9 * e_dev[0]
            diff = np.abs(s_dev - s_orc) / s_orc
            say(f"snapshot {it}: max|dp|/(max|dr| c) {ratio:.3f}, dirty {dirty.tolist()}, leaves {worst:.2e}, "
                f"tree {tree:.2e}, root sigma vs reference algorithm {gap.max():.1e}; exact spectra of the two "
                f"matrices differ by {np.abs(e_dev - e_orc).max():.3e} <= ||A_dev - A_orc||_2 = {weyl:.3e} "
                f"(sigma_0 {e_dev[0]:.1f}); root sigma vs oracle pipeline top-16 {diff[:16].max():.2e} "
                f"({time.time() - t0:.0f}s)")
    finally:
        pipe.close()


# ------------------------------------------------------- configurations 2, 3
def column_mass_bound(n, rowptr, col, rtol=1e-9):
    """c_t = alpha nd_t sum_k ((1-alpha) P^T)^k 1, summed until the mass left
    is below rtol of the total, plus that remaining mass (every later term of
    any entry is at most it), so the result is an upper bound of c."""
    from helpers import transition
    P, deg = transition(n, rowptr, col)
    nd = (deg > 0).astype(np.float64)
    PT = P.T.tocsr()
    c = np.zeros(n)
    x = np.ones(n)
    tot = float(n)
    while x.sum() > rtol * tot:
        c += x
        x = (1 - ALPHA) * (PT @ x)
    return ALPHA * nd * c + x.sum()


def sampled_push_check(pipe, opp, w, idx, dirs, fp32):
    """Sampled sources of the device against the oracle FIFO push on the same
    snapshot stream (the invariant bound), termination and conservation."""
    worst = 0.0
    for d in dirs:
        rp, col = pipe.graph.csr(d)
        deg = np.diff(rp)
        # undirected graphs have c_t <= deg(t); directed ones need the sum
        c = deg.astype(np.float64) if not w.directed else column_mass_bound(w.n, rp, col)
        op, orr = opp.p(d), opp.r(d)
        for a, i in enumerate(idx):
            p, r = pipe.source_state(int(i), d)
            assert (r[deg > 0] <= w.rmax * (1 + 1e-6)).all(), (d, i)
            assert (p >= 0).all()
            assert abs(p.sum() + r.sum() - 1.0) < (1e-4 if fp32 else 1e-9), (d, i, p.sum() + r.sum())
            dr = np.abs(orr[a] - r).max()
            slack = 2e-6 * np.abs(op[a]).max() if fp32 else 0.0
            excess = np.abs(p - op[a]) - (dr * c + slack)
            assert excess.max() <= 1e-12, (d, i, excess.max())
            worst = max(worst, (np.abs(p - op[a]) / (dr * c + slack + 1e-15)).max())
    return worst


def config_check(capi, name, nsample=32, nleaves=4):
    t0 = time.time()
    w = gen.workload(name)
    seed = 7
    say(f"{name}: n={w.n} |S|={w.subset.size} blocks={w.nparts} rank={w.rank} (gen {time.time() - t0:.0f}s)")
    pipe = pipeline(capi, w, seed)
    idx = np.linspace(0, w.subset.size - 1, nsample).astype(int)
    og = orc.Graph(w.n, w.directed)
    opp = orc.PPR(w.n, w.subset[idx], w.directed, w.alpha, w.rmax, fp32=True)
    dirs = (0, 1) if w.directed else (0,)
    starts, widths = orc.block_map(w.n, w.nparts)
    rng = np.random.Generator(np.random.PCG64(3))
    try:
        for it, (s, d) in enumerate(w.snapshots[:3], start=1):
            st = pipe.snapshot(s, d)
            og.append(s, d)
            if it == 1:
                opp.init_static()
            else:
                opp.refresh_dynamic(og, 16)
            opp.push(og, 16)
            ratio = sampled_push_check(pipe, opp, w, idx, dirs, True)
            dirty = np.flatnonzero(pipe.block_norms()[3])
            assert st["ndirty"] == dirty.size
            pick = dirty if dirty.size <= nleaves else np.unique(np.concatenate(
                [rng.choice(dirty, nleaves - 1, replace=False), [dirty[-1]]]))
            worst = 0.0
            for j in pick:
                err = leaf_error(pipe, j, starts[j], widths[j], seed, w.rank, w.q, True)
                assert err <= 2e-3, f"leaf {j} at snapshot {it}: {err:.3e}"
                worst = max(worst, err)
            # bf16x3 merge products (the precision configuration) carry ~16 bits
            tree_tol = 1e-3 if w.merge_bf16x3 else 2e-4
            tree = check_tree(pipe, w.nparts, w.fanout, w.rank, w.transform, tree_tol) if it in (1, 3) else None
            say(f"snapshot {it}: push bound ratio {ratio:.3f} over {nsample} sources; {dirty.size} dirty, leaves "
                f"{list(pick)} worst {worst:.2e}; tree {tree if tree is None else f'{tree:.2e}'} "
                f"({time.time() - t0:.0f}s)")
    finally:
        pipe.close()


def test_config_batched_sampled(capi):
    """2,097,152-node R-MAT, 16M edges, 3000 sources, rmax 1e-7, 256 blocks,
    rank 128, fan-out 4, FP32 state; initial snapshot + 2 insertions."""
    config_check(capi, "batched")


def test_config_precision_sampled(capi):
    """2,738,013-node directed Chung-Lu, 3000 sources, 512 blocks (ragged
    fan-out-4 tree), rank 128, FP32 state; initial snapshot + 2 insertions."""
    config_check(capi, "precision")


# ------------------------------------------------- configurations 4, 5 (share)
@pytest.mark.parametrize("name,shards", [("served", 3), ("atscale", 22)])
def test_config_share_properties(capi, name, shards):
    """GPU 0's share of a source-sharded run (the full |S| = 3000 state of
    these graphs exceeds one GPU: 298 GB dense for the served graph, 2 TB at
    41.6M nodes; 22 shards keep |S| = 137 >= rank + 5 at 41.6M nodes): the
    graph drawn in HBM, initial snapshot + 2 insertions,
    LP transform.  Sampled sources: termination, p >= 0, Sum p + Sum r = 1;
    sampled dirty leaves against the oracle's frPCA; sampled merges and the
    root against the FP64 spectrum of their children."""
    t0 = time.time()
    w, owned = gen.device_workload(capi, name)
    w.subset = np.array_split(w.subset, shards)[0].copy()
    say(f"{name}: n={w.n} edges={sum(x.size for x, _ in w.snapshots)} share |S|={w.subset.size} "
        f"blocks={w.nparts} (drawn in {time.time() - t0:.0f}s)")
    seed = 7
    pipe = pipeline(capi, w, seed)
    starts, widths = orc.block_map(w.n, w.nparts)
    rng = np.random.Generator(np.random.PCG64(5))
    try:
        for it, (s, d) in enumerate(w.snapshots[:3], start=1):
            st = pipe.snapshot(s, d)
            degs = [pipe.graph.degrees(dd)[0] for dd in (0, 1)]
            worst_mass = 0.0
            for i in np.linspace(0, w.subset.size - 1, 6).astype(int):
                for dd in (0, 1):
                    p, r = pipe.source_state(int(i), dd)
                    assert (r[degs[dd] > 0] <= w.rmax * (1 + 1e-6)).all(), (it, i, dd)
                    assert (p >= 0).all(), (it, i, dd)
                    worst_mass = max(worst_mass, abs(p.sum() + r.sum() - 1.0))
            assert worst_mass < 1e-4, worst_mass
            dirty = np.flatnonzero(pipe.block_norms()[3])
            assert st["ndirty"] == dirty.size
            pick = dirty if dirty.size <= 2 else rng.choice(dirty, 2, replace=False)
            worst = 0.0
            for j in pick:
                err = leaf_error(pipe, j, starts[j], widths[j], seed, w.rank, w.q, True)
                assert err <= 2e-3, f"leaf {j} at snapshot {it}: {err:.3e}"
                worst = max(worst, err)
            tree = check_tree(pipe, w.nparts, w.fanout, w.rank, w.transform, 2e-4, per_level=12, seed=it) \
                if it in (1, 3) else None
            say(f"snapshot {it}: {st['ms_total']:.0f} ms, mass error {worst_mass:.1e}, {dirty.size} dirty, leaves "
                f"{list(pick)} worst {worst:.2e}, tree {tree if tree is None else f'{tree:.2e}'} "
                f"({time.time() - t0:.0f}s)")
    finally:
        pipe.close()
        for x in owned:
            x.free()
