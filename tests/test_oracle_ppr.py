"""The oracle's refresh + push pinned by closed forms: the exact PPR matrix
(dense linear solve), termination, mass conservation, the Zhang et al.
insertion identity, and a hand-computed first push step."""
import numpy as np
import pytest

from oracle import oracle as orc
from helpers import random_graph, reserve_mass

ALPHA = 0.2


def sources_with_edges(g, k, seed, d=0):
    deg, _ = g.degrees(d)
    cand = np.flatnonzero(deg > 0)
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.choice(cand, size=min(k, cand.size), replace=False).astype(np.int32)


def invariant_residual(ppr, g, d, directed):
    """max_s | pi_s - p_s - sum_v r_s(v) pi_v | with the exact FP64 PPR."""
    rp, col = g.csr(d)
    n = g.n
    Pi, _ = orc.exact_ppr(n, rp, col, ALPHA, np.arange(n), directed)
    p, r = ppr.p(d), ppr.r(d)
    exact = Pi[ppr.sources]
    return np.abs(exact - p - r @ Pi).max()


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_static_push_invariant_and_termination(directed, seed):
    n = 40
    u, v = random_graph(n, 120, seed, directed)
    g = orc.Graph(n, directed); g.append(u, v)
    src = sources_with_edges(g, 6, seed)
    ppr = orc.PPR(n, src, directed, ALPHA, 1e-4, fp32=False)
    ppr.init_static(); ppr.push(g)
    for d in ppr.dirs():
        deg, _ = g.degrees(d)
        r = ppr.r(d)
        assert (r[:, deg > 0] <= 1e-4 + 1e-18).all()
        assert invariant_residual(ppr, g, d, directed) < 1e-12
    # the matrix accumulator equals the sum of reserves (no integer ratio yet)
    M = ppr.M()
    tot = sum(ppr.p(d) for d in ppr.dirs())
    assert np.abs(M - tot).max() < 1e-12


@pytest.mark.parametrize("directed", [False, True])
def test_dynamic_update_keeps_invariant(directed):
    n = 40
    u, v = random_graph(n, 160, 5, directed)
    g = orc.Graph(n, directed); g.append(u[:100], v[:100])
    src = sources_with_edges(g, 5, 5)
    ppr = orc.PPR(n, src, directed, ALPHA, 1e-4, fp32=False)
    ppr.init_static(); ppr.push(g)
    g.append(u[100:], v[100:])
    ppr.refresh_dynamic(g)
    # Zhang et al.: the exact-ratio update alone restores the invariant on G'
    for d in ppr.dirs():
        assert invariant_residual(ppr, g, d, directed) < 1e-12
    ppr.push(g)
    for d in ppr.dirs():
        assert invariant_residual(ppr, g, d, directed) < 1e-12


def test_integer_ratio_quirk_breaks_invariant():
    """The reference's (j+1)/j integer division (ppr:2102) is reproduced on
    request and measurably breaks the invariant (a planted-mistake check that
    the invariant test can see such a change)."""
    n = 40
    u, v = random_graph(n, 200, 9, False)
    g = orc.Graph(n, False); g.append(u[:120], v[:120])
    src = sources_with_edges(g, 5, 9)
    ppr = orc.PPR(n, src, False, ALPHA, 1e-4, fp32=False, quirk_int_ratio=True)
    ppr.init_static(); ppr.push(g)
    g.append(u[120:], v[120:])
    ppr.refresh_dynamic(g)
    assert invariant_residual(ppr, g, 0, False) > 1e-6


def test_first_push_step_by_hand():
    """Path 0-1-2 (undirected), source 0, rmax large enough to stop after the
    first push: p0 = alpha, r1 = (1-alpha), r0 = 0."""
    g = orc.Graph(3, False)
    g.append(np.array([0, 1], np.int32), np.array([1, 2], np.int32))
    ppr = orc.PPR(3, np.array([0], np.int32), False, ALPHA, 0.9, fp32=False)
    ppr.init_static(); ppr.push(g)
    assert np.allclose(ppr.p()[0], [ALPHA, 0, 0])
    assert np.allclose(ppr.r()[0], [0, 1 - ALPHA, 0])


def test_mass_conservation_larger_graph():
    n = 3000
    u, v = random_graph(n, 12000, 3, True)
    g = orc.Graph(n, True); g.append(u, v)
    src = sources_with_edges(g, 8, 3)
    ppr = orc.PPR(n, src, True, ALPHA, 1e-6, fp32=False)
    ppr.init_static(); ppr.push(g, nthreads=4)
    for d in (0, 1):
        rp, col = g.csr(d)
        mv = reserve_mass(n, rp, col, ALPHA)
        p, r = ppr.p(d), ppr.r(d)
        lhs = p.sum(1) + r @ mv
        assert np.abs(lhs - mv[src]).max() < 1e-10


def _updated_state(n, u, v, cut, directed, rmax, two_sided, groups):
    """Static push on the first `cut` edges, then the rest inserted in
    `groups` consecutive batches (update + push after each)."""
    g = orc.Graph(n, directed); g.append(u[:cut], v[:cut])
    src = sources_with_edges(g, 6, 11)
    ppr = orc.PPR(n, src, directed, ALPHA, rmax, fp32=False, two_sided=two_sided)
    ppr.init_static(); ppr.push(g)
    for part in np.array_split(np.arange(cut, u.size), groups):
        g.append(u[part], v[part])
        ppr.refresh_dynamic(g)
        ppr.push(g)
    return g, ppr


@pytest.mark.parametrize("directed", [False, True])
def test_two_sided_push_drains_negative_residues(directed):
    """The insertion update lowers r at the updated row; the reference pushes
    only r > rmax (ppr:1445), so those negative residues stay and the reserves
    drift from the exact PPR by sum_v r(v) pi_v.  Two-sided pushing (|r| >
    rmax) drains them: |r| <= rmax everywhere and, undirected, |p - pi| <=
    rmax * deg entrywise."""
    n, rmax = 60, 1e-6
    u, v = random_graph(n, 300, 12, directed)
    for two in (False, True):
        g, ppr = _updated_state(n, u, v, 150, directed, rmax, two, 3)
        for d in ppr.dirs():
            assert invariant_residual(ppr, g, d, directed) < 1e-12
            deg, _ = g.degrees(d)
            r = ppr.r(d)[:, deg > 0]
            rp, col = g.csr(d)
            pi, _ = orc.exact_ppr(n, rp, col, ALPHA, ppr.sources, directed)
            err = np.abs(ppr.p(d) - pi)
            if two:
                assert np.abs(r).max() <= rmax * (1 + 1e-12)
                if not directed:
                    assert (err <= rmax * np.maximum(deg, 1)[None, :] + 1e-15).all()
            else:
                assert r.min() < -100 * rmax          # left in place by the positive-side push
                assert err.max() > 1e3 * rmax * deg.max()


def test_two_sided_state_does_not_depend_on_batch_grouping():
    """Inserting the same edges in one batch or in several gives reserves
    within 2 rmax deg of each other with two-sided pushing; with the
    reference's positive-side push the grouping leaves a much larger gap."""
    n, rmax = 80, 1e-7
    u, v = random_graph(n, 400, 13, False)
    gaps = {}
    for two in (False, True):
        _, a = _updated_state(n, u, v, 200, False, rmax, two, 1)
        g, b = _updated_state(n, u, v, 200, False, rmax, two, 5)
        deg, _ = g.degrees(0)
        gaps[two] = (np.abs(a.p(0) - b.p(0)) / (rmax * np.maximum(deg, 1))[None, :]).max()
    assert gaps[True] <= 2.0
    assert gaps[False] > 1e3
