"""The push autotuner (tsvd_pipeline_snapshot_tuned): the push of one
snapshot is timed under each candidate setting from the same saved state, the
fastest is kept, and the snapshot ends as a plain one with that setting would
(reserves within the push bound, since push order may differ); a static
snapshot is not tuned."""
import numpy as np
import pytest

from treesvd_amd import gen
from helpers import column_mass

pytestmark = pytest.mark.gpu
CANDS = [(1024, 0.5, 1), (4096, 0.25, 8), (64, 1.0, 4)]


def workload():
    s, d, n = gen.rmat_edges(12, 24000, 6)
    snaps = gen.shuffle_split(s, d, 6, 0.7, [2400] * 2)
    sub = gen.sample_subset(snaps[0][0], 80, 6)
    return n, snaps, sub


def make(capi, n, sub):
    return capi.Pipeline(n, False, sub, 0.2, 1e-7, 8, 2, 10, 6, precision=capi.TSVD_FP64, seed=3)


def test_tuned_snapshot_keeps_a_candidate_and_matches_the_plain_push(capi):
    n, snaps, sub = workload()
    a, b = make(capi, n, sub), make(capi, n, sub)
    a.snapshot(*snaps[0])
    b.snapshot(*snaps[0])
    st, chosen, ms = a.snapshot_tuned(*snaps[1], CANDS)
    assert chosen in CANDS and (ms > 0).all()
    assert chosen == CANDS[int(np.argmin(ms))]
    assert a.ppr.push_setting == chosen
    b.ppr.push_setting = chosen
    st_b = b.snapshot(*snaps[1])
    assert st["push_pops"] > 0 and st_b["push_pops"] > 0
    rp, col = a.graph.csr(0)
    c = column_mass(n, rp, col, 0.2)
    pa, pb, ra, rb = a.ppr.p(), b.ppr.p(), a.ppr.r(), b.ppr.r()
    bound = (np.abs(ra).max() + np.abs(rb).max()) * c + 1e-12
    assert (np.abs(pa - pb) <= bound[None, :]).all()
    sa, sb = a.root_sigma(), b.root_sigma()
    assert np.abs(sa - sb).max() / sa.max() < 1e-3
    # the setting stays for the next snapshot
    a.snapshot(*snaps[2])
    assert a.ppr.push_setting == chosen


def test_static_snapshot_is_not_tuned(capi):
    n, snaps, sub = workload()
    p = make(capi, n, sub)
    before = p.ppr.push_setting             # before any push: the fallback chunk
    st, chosen, ms = p.snapshot_tuned(*snaps[0], CANDS)
    # untimed; the handle keeps its own setting, whose chunk the first push
    # chose without a caller setting (64 neighbours: this graph's hub buffers
    # are small)
    assert (ms == -1).all() and chosen == p.ppr.push_setting
    assert chosen[1:] == before[1:] and before[0] == 256 and chosen[0] == 64
