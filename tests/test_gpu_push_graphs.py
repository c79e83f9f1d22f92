"""Push rounds replayed from captured hipGraphs, through the C ABI.

With graphs on (tsvd_ppr_set_graphs; off by default) the push enqueues its
rounds in blocks of six (round kernel + split-row kernel each) captured once
per launch shape and replayed until the device's frontier count is zero.  The replay must compute what rounds launched one by
one compute (ppr:1394-1498 restated as synchronous rounds), so both forms are
checked against the oracle's FIFO push with the bound of test_gpu_push.py,
|p_gpu(t) - p_oracle(t)| <= (max|r_gpu| + max|r_oracle|) * c_t, over an
initial snapshot and insertion updates, and the graph path must actually run:
replays counted, captures reused across blocks, no failed capture.  With
profiling on, the kernels inside the graph are timed by its event nodes and
counted as launches like the others.
"""

import numpy as np
import pytest

from oracle import oracle as orc
from helpers import random_graph, reserve_mass, column_mass

pytestmark = pytest.mark.gpu
ALPHA = 0.2


def sources_of(deg, k, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.choice(np.flatnonzero(deg > 0), size=k, replace=False).astype(np.int32)


@pytest.fixture(autouse=True)
def push_rounds_only(monkeypatch):
    # pull rounds are launched one by one by design; keep them off here
    monkeypatch.delenv("TSVD_PULL_FRAC", raising=False)
    monkeypatch.delenv("TSVD_PUSH_GRAPHS", raising=False)


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("prec", [0, 1])
def test_replayed_rounds_match_the_oracle_and_the_one_by_one_rounds(capi, directed, prec):
    n = 5000
    u, v = random_graph(n, 50000, 21, directed)
    cuts = [0, 35000, 40000, 45000, u.size]
    rmax = 1e-6
    handles = {}
    for mode in ("graph", "eager"):
        g = capi.Graph(n, directed)
        g.append(u[:cuts[1]], v[:cuts[1]])
        deg, _ = g.degrees(0)
        src = sources_of(deg, 100, 21)
        ppr = capi.PPR(g, src, ALPHA, rmax, precision=prec)
        ppr.set_graphs(mode == "graph")
        handles[mode] = (g, ppr)
    og = orc.Graph(n, directed); og.append(u[:cuts[1]], v[:cuts[1]])
    opp = orc.PPR(n, src, directed, ALPHA, rmax, fp32=prec == 0)
    dirs = (0, 1) if directed else (0,)
    for k in range(len(cuts) - 1):
        if k > 0:
            og.append(u[cuts[k]:cuts[k + 1]], v[cuts[k]:cuts[k + 1]])
            opp.refresh_dynamic(og, 4)
        else:
            opp.init_static()
        opp.push(og, 4)
        for mode, (g, ppr) in handles.items():
            if k > 0:
                g.append(u[cuts[k]:cuts[k + 1]], v[cuts[k]:cuts[k + 1]])
            for d in dirs:
                ppr.refresh(capi.TSVD_REFRESH_INIT if k == 0 else capi.TSVD_REFRESH_DYNAMIC, d)
                ppr.push(d)
        for d in dirs:
            g, ppr = handles["graph"]
            rp, col = g.csr(d)
            dg = np.diff(rp)
            c = column_mass(n, rp, col, ALPHA)
            mv = reserve_mass(n, rp, col, ALPHA)
            op, orr = opp.p(d), opp.r(d)
            fp = 2e-6 * np.abs(op).max() if prec == 0 else 1e-12
            for mode, (_, q) in handles.items():
                p, r = q.p(d), q.r(d)
                assert (r[:, dg > 0] <= rmax * (1 + 1e-6)).all(), mode
                assert np.abs(p.sum(1) + r @ mv - mv[src]).max() < (1e-4 if prec == 0 else 1e-9), mode
                bound = (np.abs(r).max() + np.abs(orr).max()) * c + fp
                assert (np.abs(p - op) <= bound[None, :]).all(), (mode, k, d)
    launches, captures, failures = handles["graph"][1].graph_stats()
    print(f"graph replays {launches}, captures {captures}, failures {failures}")
    assert failures == 0
    assert launches > captures > 0, "captured blocks are replayed"
    assert handles["eager"][1].graph_stats() == (0, 0, 0)


def test_profiled_replays_time_every_kernel(capi):
    """Kernels inside a replayed graph are timed by its event nodes: six
    round and six split-row launches per replay, each with a time."""
    n = 4000
    u, v = random_graph(n, 40000, 23, False)
    g = capi.Graph(n, False); g.append(u, v)
    deg, _ = g.degrees(0)
    ppr = capi.PPR(g, sources_of(deg, 64, 23), ALPHA, 1e-6)
    ppr.set_graphs(True)
    capi.profile_reset()
    capi.profile_enable(True)
    try:
        ppr.refresh(capi.TSVD_REFRESH_INIT)
        ppr.push()
    finally:
        capi.profile_enable(False)
    prof = capi.profile_read()
    launches, captures, failures = ppr.graph_stats()
    assert failures == 0 and launches > 0
    rounds, hubs = prof["ppr_push_round"], prof["ppr_push_hub"]
    assert rounds[0] == 6 * launches and hubs[0] == 6 * launches, (prof, launches)
    assert rounds[1] > 0.0 and hubs[1] > 0.0
    st = ppr.last_push_stats(0)
    assert 0 < st["rounds"] <= 6 * launches
