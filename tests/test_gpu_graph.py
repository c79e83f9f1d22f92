"""Device CSR ingest through the C ABI: bit-exact rows against the golden
rows the reference's own Graph_dynamic.h wrote, and against the oracle over
multi-snapshot random graphs (with and without CSR rebuilds)."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc
from helpers import random_graph

pytestmark = pytest.mark.gpu
GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ingest_*.json")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_gpu_ingest_matches_reference_rows(capi, path, tmp_path):
    with open(path) as f:
        fx = json.load(f)
    n, directed = fx["n"], fx["directed"]
    g = capi.Graph(n, directed, slack=0.1)
    for k, (text, snap) in enumerate(zip(fx["files"], fx["snapshots"])):
        p = tmp_path / f"s{k}.txt"
        p.write_text(text)
        s, d = capi.read_edge_file(str(p), dup_last=True)
        g.append(s, d)
        assert g.num_edges() == snap["m"]
        for dd in range(2 if directed else 1):
            deg, former = g.degrees(dd)
            rp, col = g.csr(dd)
            for v in range(n):
                gdeg, gformer, grow = snap["rows"][dd][v]
                assert deg[v] == gdeg and former[v] == gformer, (k, dd, v)
                assert col[rp[v]:rp[v + 1]].tolist() == grow, (k, dd, v)


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("slack", [0.0, 0.1, 1.0])
def test_gpu_ingest_matches_oracle(capi, directed, slack):
    n = 5000
    u, v = random_graph(n, 60000, 21, directed, no_self_loops=False)
    cuts = [0, 30000, 30001, 40000, 52000, u.size]
    g = capi.Graph(n, directed, slack=slack)
    o = orc.Graph(n, directed)
    for a, b in zip(cuts[:-1], cuts[1:]):
        g.append(u[a:b], v[a:b])
        o.append(u[a:b], v[a:b])
        for d in range(2 if directed else 1):
            gd, gf = g.degrees(d)
            od, of = o.degrees(d)
            assert np.array_equal(gd, od) and np.array_equal(gf, of)
            grp, gcol = g.csr(d)
            orp, ocol = o.csr(d)
            assert np.array_equal(grp, orp) and np.array_equal(gcol, ocol)


def test_gpu_ingest_empty_batch_and_bad_ids(capi):
    g = capi.Graph(10, False)
    g.append(np.array([1, 2], np.int32), np.array([3, 4], np.int32))
    g.append(np.zeros(0, np.int32), np.zeros(0, np.int32))
    deg, former = g.degrees()
    assert deg.sum() == 4 and np.array_equal(deg, former)
    with pytest.raises(capi.TreeSVDError):
        g.append(np.array([0], np.int32), np.array([10], np.int32))
    with pytest.raises(capi.TreeSVDError):
        g.append(np.array([-1], np.int32), np.array([1], np.int32))
    # a rejected batch leaves the graph unchanged
    assert g.num_edges() == 2


def test_gpu_ingest_device_pointers(capi):
    n = 1000
    u, v = random_graph(n, 5000, 4, False)
    g = capi.Graph(n, False)
    du, dv = capi.DeviceArray(u.astype(np.int32)), capi.DeviceArray(v.astype(np.int32))
    assert np.array_equal(du.numpy(), u)
    g.append(du, dv)
    o = orc.Graph(n, False); o.append(u, v)
    assert np.array_equal(g.csr()[1], o.csr()[1])
