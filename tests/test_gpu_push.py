"""Device forward push and dynamic update through the C ABI.

Push order differs from the reference's FIFO, so reserves are compared with
bounds that follow from the push invariant pi_s = p_s + sum_v r_s(v) pi_v:
  * tiny graphs: the invariant itself against the exact FP64 PPR matrix;
  * larger graphs: mass conservation (sum_t p_s(t) + sum_v r_s(v) m_v = m_s,
    m = reserve mass) and the entrywise bound
    |p_gpu(t) - p_oracle(t)| <= (max|r_gpu| + max|r_oracle|) * c_t,
    c_t = sum_v pi_v(t) (c_t <= deg(t) on undirected graphs);
  * termination: every vertex with degree > 0 ends with r <= rmax.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from helpers import random_graph, reserve_mass, column_mass

pytestmark = pytest.mark.gpu
ALPHA = 0.2


@pytest.fixture(params=["hybrid", "push", "pull"], autouse=True)
def round_mode(request, monkeypatch):
    """Every push test runs with the library's hybrid schedule, with push
    rounds only, and with pull rounds only (TSVD_PULL_FRAC)."""
    frac = {"hybrid": None, "push": "0", "pull": "1e-9"}[request.param]
    if frac is None:
        monkeypatch.delenv("TSVD_PULL_FRAC", raising=False)
    else:
        monkeypatch.setenv("TSVD_PULL_FRAC", frac)
    return request.param


def pick_sources(deg, k, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.choice(np.flatnonzero(deg > 0), size=k, replace=False).astype(np.int32)


def dirs(directed):
    return (0, 1) if directed else (0,)


@pytest.mark.parametrize("two", [False, True], ids=["one-sided", "two-sided"])
@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("prec", [0, 1])
def test_gpu_push_tiny_exact_invariant(capi, directed, prec, two):
    n = 48
    u, v = random_graph(n, 150, 7, directed)
    g = capi.Graph(n, directed); g.append(u[:110], v[:110])
    deg, _ = g.degrees(0)
    src = pick_sources(deg, 10, 7)
    rmax = 1e-5
    ppr = capi.PPR(g, src, ALPHA, rmax, precision=prec)
    ppr.set_two_sided(two)
    tol = 1e-12 if prec == 1 else 2e-6
    for snap in range(2):
        if snap == 1:
            g.append(u[110:], v[110:])
        for d in dirs(directed):
            ppr.refresh(capi.TSVD_REFRESH_INIT if snap == 0 else capi.TSVD_REFRESH_DYNAMIC, d)
            ppr.push(d)
            rp, col = g.csr(d)
            Pi, _ = orc.exact_ppr(n, rp, col, ALPHA, np.arange(n), directed)
            p, r = ppr.p(d), ppr.r(d)
            dd = np.diff(rp)
            assert (r[:, dd > 0] <= rmax * (1 + 1e-6)).all()
            if two:
                assert (r[:, dd > 0] >= -rmax * (1 + 1e-6)).all()
            assert np.abs(Pi[src] - p - r @ Pi).max() < tol


@pytest.mark.parametrize("two", [False, True], ids=["one-sided", "two-sided"])
@pytest.mark.parametrize("directed", [False, True])
def test_gpu_push_matches_oracle_multi_snapshot(capi, directed, two):
    """With the reference's positive-side threshold the insertion update leaves
    negative residues, so the entrywise bound is loose; two-sided pushing
    keeps |r| <= rmax and the bound tight (|p_gpu - p_oracle| <= 2 rmax c)."""
    n = 4000
    u, v = random_graph(n, 40000, 11, directed)
    cuts = [0, 28000, 32000, 36000, u.size]
    rmax = 1e-6
    g = capi.Graph(n, directed)
    g.append(u[:cuts[1]], v[:cuts[1]])
    deg, _ = g.degrees(0)
    src = pick_sources(deg, 70, 11)        # two source groups, one partial
    ppr = capi.PPR(g, src, ALPHA, rmax, precision=capi.TSVD_FP64)
    ppr.set_two_sided(two)
    og = orc.Graph(n, directed); og.append(u[:cuts[1]], v[:cuts[1]])
    opp = orc.PPR(n, src, directed, ALPHA, rmax, fp32=False, two_sided=two)
    for k in range(len(cuts) - 1):
        if k > 0:
            g.append(u[cuts[k]:cuts[k + 1]], v[cuts[k]:cuts[k + 1]])
            og.append(u[cuts[k]:cuts[k + 1]], v[cuts[k]:cuts[k + 1]])
            opp.refresh_dynamic(og, 4)
        else:
            opp.init_static()
        opp.push(og, 4)
        for d in dirs(directed):
            ppr.refresh(capi.TSVD_REFRESH_INIT if k == 0 else capi.TSVD_REFRESH_DYNAMIC, d)
            ppr.push(d)
            st = ppr.last_push_stats(d)
            assert st["rounds"] > 0 and st["pops"] > 0
            rp, col = g.csr(d)
            p, r = ppr.p(d), ppr.r(d)
            dg = np.diff(rp)
            assert (r[:, dg > 0] <= rmax * (1 + 1e-9)).all()
            if two:
                assert (r[:, dg > 0] >= -rmax * (1 + 1e-9)).all()
            mv = reserve_mass(n, rp, col, ALPHA)
            assert np.abs(p.sum(1) + r @ mv - mv[src]).max() < 1e-9
            c = column_mass(n, rp, col, ALPHA)
            op, orr = opp.p(d), opp.r(d)
            bound = (np.abs(r).max() + np.abs(orr).max()) * c + 1e-12
            assert (np.abs(p - op) <= bound[None, :]).all()
            if not directed:
                assert (c <= dg + 1e-9).all()


def test_gpu_push_fp32_state(capi):
    n = 3000
    u, v = random_graph(n, 30000, 13, False)
    g = capi.Graph(n, False); g.append(u, v)
    deg, _ = g.degrees(0)
    src = pick_sources(deg, 64, 13)
    rmax = 1e-6
    ppr = capi.PPR(g, src, ALPHA, rmax, precision=capi.TSVD_FP32)
    ppr.refresh(capi.TSVD_REFRESH_INIT); ppr.push()
    og = orc.Graph(n, False); og.append(u, v)
    opp = orc.PPR(n, src, False, ALPHA, rmax, fp32=True)
    opp.init_static(); opp.push(og, 4)
    p, r = ppr.p(), ppr.r()
    rp, col = g.csr()
    c = column_mass(n, rp, col, ALPHA)
    # FP32 storage adds rounding of order 1e-7 relative per update
    bound = (np.abs(r).max() + np.abs(opp.r()).max()) * c + 2e-6 * np.abs(opp.p()).max()
    assert (np.abs(p - opp.p()) <= bound[None, :]).all()
    mv = reserve_mass(n, rp, col, ALPHA)
    assert np.abs(p.sum(1) + r @ mv - mv[src]).max() < 1e-4


def test_gpu_push_source_range(capi):
    """Refreshing and pushing a sub-range of sources [start, end), as the
    reference's per-thread calls do, leaves the other sources untouched."""
    n = 500
    u, v = random_graph(n, 3000, 17, False)
    g = capi.Graph(n, False); g.append(u, v)
    deg, _ = g.degrees(0)
    src = pick_sources(deg, 100, 17)
    ppr = capi.PPR(g, src, ALPHA, 1e-5, precision=capi.TSVD_FP64)
    ppr.refresh(capi.TSVD_REFRESH_INIT, 0, 0, 40)
    ppr.push(0, 0, 40)
    p = ppr.p()
    assert (p[:40].sum(1) > 0).all() and (p[40:] == 0).all()
    ppr.refresh(capi.TSVD_REFRESH_INIT, 0, 40, 100)
    ppr.push(0, 40, 100)
    p2 = ppr.p()
    assert np.array_equal(p2[:40], p[:40]) and (p2[40:].sum(1) > 0).all()
