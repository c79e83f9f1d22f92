"""The served form (tsvd_server_*) through the C ABI.

A server admits every queued insertion batch into one snapshot.  The claim
checked here: each request's result is the same in the mix as alone.  The
result handed to a request is the embedding of the state that includes every
request up to its step's last ticket.  The plain method (one snapshot per
request, applied one by one) must give the same embedding for that same
prefix, within a stated tolerance.

What separates the two is only how insertions are grouped into snapshots.
The lazy test is off here (delta = 0, every block refactored) and sketches
are per block, so the grouping acts through the reserves alone:
  * two-sided pushing (|r| > rmax) keeps every reserve within rmax * deg of the
    exact PPR whatever the grouping: root sigma within 1e-2 relative (2e-3 for
    the top half) and the top-half subspace within 1e-5 at rmax = 1e-7, bars
    that cover the push-order spread of two identical runs (see TOL);
  * the reference's positive-side threshold leaves the insertion update's
    negative residues in place, and those depend on the grouping: root sigma
    within 6e-2 (2.5e-2 top half), top-half subspace within 1.5e-2.
Both bounds were sized on the oracle pipeline (same graph, grouped 1+3+4 vs
one by one), which shows 4.5e-4 / 1.1e-4 / 4e-7 and 3.3e-2 / 6.8e-3 / 3.4e-3;
the device, grouped 1+2+6, shows 3.7e-4 / 1.9e-4 / 3.5e-7 and 3.1e-2 /
1.1e-2 / 5.1e-3."""
import threading

import numpy as np
import pytest

from treesvd_amd import gen

pytestmark = pytest.mark.gpu


def workload(seed=4):
    s, d, n = gen.rmat_edges(12, 30000, seed)
    snaps = gen.shuffle_split(s, d, seed, 0.6, [1500, 400, 2500, 900, 3000, 1200, 600, 2000])
    sub = gen.sample_subset(snaps[0][0], 96, seed)
    return n, snaps, sub


ARGS = dict(alpha=0.2, rmax=1e-7, nparts=8, fanout=2, rank=12, q=6, precision=1, delta=0.0, seed=5,
            two_sided=True)
# sigma all, sigma top half, top-half subspace.  Two-sided: two runs of the
# same pipeline on the same stream already differ by their push order (float
# atomics), which the NC transform log10(x / rmax) amplifies for reserves near
# rmax; their measured spread is up to 5e-4 on the top half, 4e-3 on the last
# sigma at the rank-12 cut and 8e-7 on the subspace, and the bars leave 2-4x
# of margin over it.
TOL = {True: (1e-2, 2e-3, 1e-5), False: (6e-2, 2.5e-2, 1.5e-2)}


def sigma_of(E):
    """NC root E = U sqrt(S): the column norms squared are S."""
    return np.sort((E ** 2).sum(axis=0))[::-1]


def same_embedding(Ea, Eb, k, two_sided=True, what=""):
    sa, sb = sigma_of(Ea), sigma_of(Eb)
    rel = np.abs(sa - sb) / sa
    Ua, _ = np.linalg.qr(Ea)
    Ub, _ = np.linalg.qr(Eb)
    h = k // 2
    sub = 1.0 - np.linalg.norm(Ua[:, :h].T @ Ub) ** 2 / h
    t_all, t_half, t_sub = TOL[two_sided]
    msg = f"{what}: sigma rel {rel.max():.2e} (top half {rel[:h].max():.2e}), subspace {sub:.2e}"
    print(msg)
    assert rel.max() < t_all and rel[:h].max() < t_half and sub < t_sub, msg


def plain_prefix_embeddings(capi, n, snaps, sub, **kw):
    """Embedding after each request applied alone (one snapshot each)."""
    pipe = capi.Pipeline(n, False, sub, **{**ARGS, **kw})
    out = []
    for s, d in snaps:
        pipe.snapshot(s, d)
        out.append(pipe.embedding())
    pipe.close()
    return out


@pytest.mark.parametrize("two", [True, False], ids=["two-sided", "reference-threshold"])
def test_served_results_equal_plain_prefix(capi, two):
    n, snaps, sub = workload()
    alone = plain_prefix_embeddings(capi, n, snaps, sub, two_sided=two)
    srv = capi.Server(n, False, sub, **{**ARGS, "two_sided": two})
    t0 = srv.submit(*snaps[0])
    E0, r0 = srv.wait(t0)
    assert r0["step_requests"] == 1 and r0["last_ticket"] == 0
    # the stream: everything queued at once, so steps coalesce several requests
    tickets = [srv.submit(s, d) for s, d in snaps[1:]]
    results = [srv.wait(t) for t in tickets]
    steps = sorted({r["snapshot"] for _, r in results})
    assert len(steps) < len(tickets), "requests queued together must share steps"
    for (E, r), t in zip(results, tickets):
        assert r["ticket"] == t and r["last_ticket"] >= t
        assert r["latency_ms"] >= r["done_ms"] - r["start_ms"] - 1e-6
        same_embedding(alone[r["last_ticket"]], E, ARGS["rank"], two,
                       f"ticket {t} (state {r['last_ticket']}, step of {r['step_requests']})")
    # requests of one step all get that step's embedding and stats
    by_step = {}
    for E, r in results:
        by_step.setdefault(r["snapshot"], []).append((E, r))
    for group in by_step.values():
        assert all(np.array_equal(group[0][0], E) for E, _ in group)
        assert all(r["step_requests"] == len(group) for _, r in group)
        assert sum(np.asarray(snaps[r["ticket"]][0]).size for _, r in group) == group[0][1]["step_edges"]
    srv.close()


def test_served_request_cap_one_is_the_plain_method(capi):
    n, snaps, sub = workload(seed=7)
    alone = plain_prefix_embeddings(capi, n, snaps[:5], sub)
    srv = capi.Server(n, False, sub, max_step_requests=1, **ARGS)
    tickets = [srv.submit(s, d) for s, d in snaps[:5]]
    for t in tickets:
        E, r = srv.wait(t)
        assert r["step_requests"] == 1 and r["last_ticket"] == t and r["snapshot"] == t + 1
        same_embedding(alone[t], E, ARGS["rank"], True, f"ticket {t}")
    srv.close()


def test_served_concurrent_clients_and_edge_budget(capi):
    """Several client threads submit while steps run; an edge budget splits
    the queue into steps of at most that many edges (one oversized request
    still runs alone)."""
    n, snaps, sub = workload(seed=9)
    budget = 3000
    srv = capi.Server(n, False, sub, max_step_edges=budget, **ARGS)
    assert srv.edge_budget == budget
    srv.wait(srv.submit(*snaps[0]))
    got, lock = [], threading.Lock()

    def client(parts):
        for s, d in parts:
            t = srv.submit(s, d)
            E, r = srv.wait(t)
            with lock:
                got.append(r)

    rest = snaps[1:]
    th = [threading.Thread(target=client, args=(rest[i::3],)) for i in range(3)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert sorted(r["ticket"] for r in got) == list(range(1, len(snaps)))
    for r in got:
        assert r["step_edges"] <= budget or r["step_requests"] == 1
    srv.close()


def test_served_rejects_bad_ids_and_unknown_tickets(capi):
    n, snaps, sub = workload()
    srv = capi.Server(n, False, sub, **ARGS)
    with pytest.raises(capi.TreeSVDError):
        srv.submit(np.array([0, n], np.int32), np.array([1, 2], np.int32))
    with pytest.raises(capi.TreeSVDError):
        srv.wait(5)
    t = srv.submit(*snaps[0])
    srv.wait(t, embedding=False)
    with pytest.raises(capi.TreeSVDError):
        srv.wait(t)                      # collected once
    srv.close()


def test_served_overlap_matches_steps_one_at_a_time(capi):
    """Back parts of step t (frPCA, merges, root) overlap the front part of
    step t+1 on another stream; with the same steps (cap 1) the results equal
    those of a server that runs steps one at a time, within the push-order
    tolerance (TOL), and requests complete in order.  A second serial server
    shows the run-to-run spread the tolerance covers."""
    n, snaps, sub = workload(seed=11)
    outs = {}
    for key, overlap in (("overlap", True), ("serial", False), ("serial again", False)):
        srv = capi.Server(n, False, sub, max_step_requests=1, overlap=overlap, **ARGS)
        tickets = [srv.submit(s, d) for s, d in snaps]
        res = [srv.wait(t) for t in tickets]
        assert [r["snapshot"] for _, r in res] == list(range(1, len(snaps) + 1))
        assert all(r["done_ms"] >= r["start_ms"] for _, r in res)
        outs[key] = [E for E, _ in res]
        srv.close()
    for other in ("overlap", "serial again"):
        for t, (Ea, Eb) in enumerate(zip(outs[other], outs["serial"])):
            same_embedding(Eb, Ea, ARGS["rank"], True, f"ticket {t}, {other} vs one at a time")
