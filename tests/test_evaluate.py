"""The downstream quality harness (treesvd_amd.evaluate), CPU only: the
top-k one-vs-rest rule and the LP precision rule on hand-checked cases."""
import numpy as np
import pytest

from oracle import oracle as orc
from treesvd_amd import evaluate


def test_link_precision_matches_the_oracle_rule():
    rng = np.random.Generator(np.random.PCG64(1))
    s = rng.standard_normal(200)
    pos = rng.random(200) < 0.3
    assert evaluate.link_precision(s, pos) == orc.lp_precision(s, pos, int(pos.sum()))
    assert evaluate.link_precision([3, 2, 1], [True, False, True], 2) == 0.5


def test_node_classification_separable_and_random(tmp_path):
    # two well-separated classes in 4 dimensions: perfect F1; one node has two labels
    rng = np.random.Generator(np.random.PCG64(0))
    n = 200
    lab = rng.integers(0, 2, n)
    E = rng.standard_normal((n, 4)) * 0.1
    E[:, 0] += np.where(lab == 1, 5.0, -5.0)
    Y = [[str(c)] for c in lab]
    res = evaluate.node_classification(E, Y, 0.5, seed=0)
    assert res["micro"] == pytest.approx(1.0) and res["macro"] == pytest.approx(1.0)
    # labels unrelated to the embedding: about chance
    Y2 = [[str(c)] for c in rng.integers(0, 2, n)]
    assert 0.3 < evaluate.node_classification(E, Y2, 0.5, seed=0)["micro"] < 0.7
    p = tmp_path / "labels.txt"
    p.write_text("4 a b\n7 c\n\n9 a\n")
    X, Yr = evaluate.read_node_labels(str(p))
    assert X == [4, 7, 9] and Yr == [["a", "b"], ["c"], ["a"]]
