"""The oracle's n >= 1e7-node path (Tree_d_LP.cpp:1377-2024,
All_initialization_and_push_final at ppr:6188-6832, transform at
ppr:6898-7229), pinned without running the reference (its build needs Eigen,
emhash and MKL):

  * insertion update in batch order: where no vertex gets its first edge it
    must equal the per-row update (itself pinned by the exact PPR invariant)
    and keep the invariant pi_s = p + sum_v r(v) pi_v on the new graph;
  * Q8: when a vertex w of former degree 0 gets its first edge, the update
    drops w's residue, including what earlier edges of the batch added to it.
    The exact invariant then fails by exactly coef * pi_w, with coef = the
    dropped amount: r_w before the batch, plus the add of an edge into w
    (the textbook (1-alpha) p_x / (d_x alpha) of one insertion at x) when
    that edge comes first in the batch, and nothing for it when it comes later;
  * negative residues are dropped when the state is stored;
  * Q11: rows >= 30 floor(|S|/30) of every block stay empty, and rows are
    otherwise independent (sources do not interact), so |S| = 64 equals
    |S| = 60 on the first 60 rows.
"""
import numpy as np

from oracle import oracle as orc
from helpers import random_graph

ALPHA = 0.2


def ring_graph(n, m, seed):
    """Directed graph where every vertex has in- and out-degree >= 1."""
    u, v = random_graph(n, m, seed, True)
    ring = np.arange(n, dtype=np.int32)
    return np.concatenate([ring, u]), np.concatenate([(ring + 1) % n, v]).astype(np.int32)


def invariant_defect(g, pp, d, src, n):
    rp, col = g.csr(d)
    Pi, _ = orc.exact_ppr(n, rp, col, ALPHA, np.arange(n), True)
    return Pi[src] - pp.p(d) - pp.r(d) @ Pi, Pi


def test_edge_order_update_equals_row_update_without_first_edges():
    n = 60
    u, v = ring_graph(n, 240, 3)
    cut = 200
    src = np.array([0, 7, 19, 33], np.int32)
    states = []
    for edges in (False, True):
        g = orc.Graph(n, True)
        g.append(u[:cut], v[:cut])
        pp = orc.PPR(n, src, True, ALPHA, 1e-9, fp32=False)
        pp.init_static()
        pp.push(g)
        g.append(u[cut:], v[cut:])
        if edges:
            pp.refresh_edges(g, u[cut:], v[cut:])
        else:
            pp.refresh_dynamic(g)
        states.append([(pp.p(d).copy(), pp.r(d).copy()) for d in (0, 1)])
        for d in (0, 1):
            D, _ = invariant_defect(g, pp, d, src, n)
            assert np.abs(D).max() < 1e-12
    for d in (0, 1):
        assert np.abs(states[0][d][0] - states[1][d][0]).max() < 1e-14
        assert np.abs(states[0][d][1] - states[1][d][1]).max() < 1e-14


def q8_case(order):
    # w = 4 has no out-edge, so the static push leaves residue there; the batch
    # gives it its first out-edge (4 -> 5) and an in-edge from x = 1.
    n = 6
    u = np.array([0, 1, 2, 1, 3, 3, 5, 0], np.int32)
    v = np.array([1, 2, 0, 3, 4, 0, 0, 5], np.int32)
    batch = [(1, 4), (4, 5)] if order == "in-edge first" else [(4, 5), (1, 4)]
    bs = np.array([e[0] for e in batch], np.int32)
    bd = np.array([e[1] for e in batch], np.int32)
    src = np.array([0, 2], np.int32)
    g = orc.Graph(n, True)
    g.append(u, v)
    pp = orc.PPR(n, src, True, ALPHA, 1e-12, fp32=False)
    pp.init_static()
    pp.push(g)
    deg, _ = g.degrees(0)
    r4, p1, d1 = pp.r(0)[:, 4].copy(), pp.p(0)[:, 1].copy(), int(deg[1])
    assert (r4 > 1e-3).all() and deg[4] == 0
    g.append(bs, bd)
    pp.refresh_edges(g, bs, bd)
    D, Pi = invariant_defect(g, pp, 0, src, n)
    add = (1 - ALPHA) * p1 / (d1 * ALPHA)
    coef = r4 + add if order == "in-edge first" else r4
    return D, coef[:, None] * Pi[4][None, :], pp


def test_first_edge_of_a_dangling_vertex_drops_its_residue_in_batch_order():
    for order in ("in-edge first", "out-edge first"):
        D, expect, pp = q8_case(order)
        assert np.abs(D - expect).max() < 1e-12, order
        assert np.abs(D).max() > 1e-3, order            # the drop is material


def test_transpose_direction_keeps_the_invariant_in_the_q8_case():
    """In-degrees: every vertex the batch gives an in-edge had one before."""
    _, _, pp = q8_case("in-edge first")
    n = 6
    g = orc.Graph(n, True)
    g.append(np.array([0, 1, 2, 1, 3, 3, 5, 0, 1, 4], np.int32), np.array([1, 2, 0, 3, 4, 0, 0, 5, 4, 5], np.int32))
    D, _ = invariant_defect(g, pp, 1, np.array([0, 2], np.int32), n)
    assert np.abs(D).max() < 1e-12


def test_drop_negative_keeps_exactly_the_positive_entries():
    n = 80
    u, v = random_graph(n, 400, 5, True)
    cut = 300
    src = np.array([1, 2, 3], np.int32)
    g = orc.Graph(n, True)
    g.append(u[:cut], v[:cut])
    pp = orc.PPR(n, src, True, ALPHA, 1e-7, fp32=True)
    pp.init_static()
    pp.push(g)
    g.append(u[cut:], v[cut:])
    pp.refresh_edges(g, u[cut:], v[cut:])
    before = [(pp.p(d).copy(), pp.r(d).copy()) for d in (0, 1)]
    assert min(b[1].min() for b in before) < 0.0        # the update left negative residues
    pp.drop_negative()
    for d in (0, 1):
        p0, r0 = before[d]
        assert np.array_equal(pp.p(d), p0) and (p0 >= 0).all()
        assert np.array_equal(pp.r(d), np.where(r0 > 0, r0, 0.0))


def sparse_run(sources, n, snaps, nparts=4, seed=2):
    pipe = orc.Pipeline(n, True, sources, ALPHA, 1e-7, nparts, 2, 8, 6, mode="LP", delta=0.3,
                        dyn_start=2, lazy_start=2, seed=seed, nthreads=4, sparse_state=True)
    out = []
    for s, d in snaps:
        pipe.snapshot(s, d)
        out.append((pipe.matrix().copy(), [a.copy() for a in pipe.A_old], pipe.root_sigma.copy()))
    return pipe, out


def test_rows_past_thirty_intervals_stay_empty_and_rows_are_independent():
    n = 500
    u, v = random_graph(n, 4000, 9, True)
    snaps = [(u[:3000], v[:3000]), (u[3000:3500], v[3000:3500]), (u[3500:], v[3500:])]
    rng = np.random.Generator(np.random.PCG64(4))
    src64 = rng.choice(np.unique(u[:3000]), 64, replace=False).astype(np.int32)
    p64, o64 = sparse_run(src64, n, snaps)
    p60, o60 = sparse_run(src64[:60], n, snaps)
    assert p64.keep == 60 and p60.keep == 60
    for (M64, A64, s64), (M60, A60, s60) in zip(o64, o60):
        assert not M64[60:].any()
        assert np.array_equal(M64[:60], M60)
        for a, b in zip(A64, A60):
            assert not a[60:].any()
            assert np.array_equal(a[:60], b)
        # zero rows change only the rounding of the factorisations (frPCA's
        # LU-orthonormalised power passes round at about 1e-7 relative)
        assert np.allclose(s64, s60, rtol=1e-6, atol=1e-12)
    assert not np.abs(p64.E[60:]).any()


def test_first_snapshot_equals_the_dense_path():
    """Before any insertion nothing is dropped, and the blocks of both paths
    are log10(1 + (p + p^T)/rmax): same matrices, same root.  In FP64 state,
    where the dense path's matrix accumulator and p add the same terms (with
    FP32 state p is rounded at every store and the accumulator is not)."""
    n = 400
    u, v = random_graph(n, 3000, 12, True)
    rng = np.random.Generator(np.random.PCG64(8))
    src = rng.choice(np.unique(u), 60, replace=False).astype(np.int32)
    kw = dict(mode="LP", delta=0.3, dyn_start=2, lazy_start=2, seed=1, nthreads=4, fp32=False)
    a = orc.Pipeline(n, True, src, ALPHA, 1e-7, 4, 2, 8, 6, sparse_state=True, **kw)
    b = orc.Pipeline(n, True, src, ALPHA, 1e-7, 4, 2, 8, 6, **kw)
    a.snapshot(u, v)
    b.snapshot(u, v)
    assert a.root_sigma is not None and b.root_sigma is not None
    for x, y in zip(a.A_old, b.A_old):
        assert np.abs(x - y).max() <= 1e-12 * max(1.0, np.abs(y).max())
    # frPCA's LU-orthonormalised power passes round at about 1e-7 relative
    assert np.allclose(a.root_sigma, b.root_sigma, rtol=1e-6)


def test_edge_order_update_is_the_same_on_any_thread_count():
    """Sources are independent; each thread keeps its own per-vertex edge count."""
    n = 120
    u, v = random_graph(n, 900, 21, True)
    cut = 700
    src = np.arange(0, 40, 3, dtype=np.int32)
    out = []
    for threads in (1, 4):
        g = orc.Graph(n, True)
        g.append(u[:cut], v[:cut])
        pp = orc.PPR(n, src, True, ALPHA, 1e-7, fp32=True)
        pp.init_static()
        pp.push(g, threads)
        g.append(u[cut:], v[cut:])
        pp.refresh_edges(g, u[cut:], v[cut:], threads)
        out.append([(pp.p(d).copy(), pp.r(d).copy()) for d in (0, 1)])
    for d in (0, 1):
        assert np.array_equal(out[0][d][0], out[1][d][0]) and np.array_equal(out[0][d][1], out[1][d][1])
