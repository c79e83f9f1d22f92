"""The oracle's ingest and edge-file reader against golden rows written by
the reference's own Graph_dynamic.h (tests/golden/ingest_*.json)."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ingest_*.json")))


def load(path):
    with open(path) as f:
        return json.load(f)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_ingest_matches_reference(path, tmp_path):
    fx = load(path)
    n, directed = fx["n"], fx["directed"]
    g = orc.Graph(n, directed)
    m_total = 0
    for k, (text, snap) in enumerate(zip(fx["files"], fx["snapshots"])):
        p = tmp_path / f"s{k}.txt"
        p.write_text(text)
        s, d = orc.read_edge_file(str(p), dup_last=True)
        g.append(s, d)
        m_total += s.size
        assert m_total == snap["m"]
        for dd in range(2 if directed else 1):
            deg, former = g.degrees(dd)
            rp, col = g.csr(dd)
            for v in range(n):
                gdeg, gformer, grow = snap["rows"][dd][v]
                assert deg[v] == gdeg and former[v] == gformer
                assert col[rp[v]:rp[v + 1]].tolist() == grow


def test_reader_duplicate_last_edge(tmp_path):
    p = tmp_path / "e.txt"
    p.write_text("1 2\n3 4\n")
    s, d = orc.read_edge_file(str(p), dup_last=True)
    assert s.tolist() == [1, 3, 3] and d.tolist() == [2, 4, 4]
    s, d = orc.read_edge_file(str(p), dup_last=False)
    assert s.tolist() == [1, 3] and d.tolist() == [2, 4]
    p.write_text("1 2\n3 4")
    s, d = orc.read_edge_file(str(p), dup_last=True)
    assert s.tolist() == [1, 3]


def test_out_of_range_edge_rejected():
    g = orc.Graph(4, False)
    with pytest.raises(ValueError):
        g.append(np.array([0], np.int32), np.array([4], np.int32))
