"""Downstream quality of the subset embedding, as the reference measures it.

Node classification (labelclassification_dynamic.py:12-129): one-vs-rest
logistic regression on the embedding rows of the labelled subset, trained on
a seeded share of the nodes; every test node is assigned its k most probable
labels, k = its true label count; micro and macro F1.  Link prediction
(Tree_u_LP.cpp:1047-1110): precision among the top-|positives| scored pairs
(oracle.lp_precision is the same rule in the test oracle).
"""
from __future__ import annotations

import numpy as np


def read_node_labels(path: str):
    """Label file of labelclassification_dynamic.py:81-94: `node label...` per line."""
    X, Y = [], []
    with open(path) as f:
        for line in f:
            v = line.strip().split()
            if not v:
                continue
            X.append(int(v[0]))
            Y.append(v[1:])
    return X, Y


def node_classification(E, Y, train_percent: float, seed: int = 0):
    """Micro / macro F1 of top-k one-vs-rest logistic regression on rows of
    E (one row per labelled node, in label-file order) with label lists Y,
    trained on the first train_percent of a seeded permutation."""
    from sklearn.linear_model import LogisticRegression
    from sklearn.metrics import f1_score
    from sklearn.multiclass import OneVsRestClassifier
    from sklearn.preprocessing import MultiLabelBinarizer

    E = np.asarray(E, dtype=np.float64)
    n = len(Y)
    perm = np.random.RandomState(seed).permutation(np.arange(n))
    ntr = int(train_percent * n)
    tr, te = perm[:ntr], perm[ntr:]
    binz = MultiLabelBinarizer(sparse_output=True)
    binz.fit(Y)
    clf = OneVsRestClassifier(LogisticRegression())
    clf.fit(E[tr], binz.transform([Y[i] for i in tr]))
    probs = np.asarray(clf.predict_proba(E[te]))
    pred = np.zeros_like(probs)
    for r, i in enumerate(te):
        top = probs[r].argsort()[-len(Y[i]):]
        pred[r, top] = 1
    truth = binz.transform([Y[i] for i in te])
    return {"micro": float(f1_score(truth, pred, average="micro")),
            "macro": float(f1_score(truth, pred, average="macro"))}


def link_precision(scores, positive, m=None):
    """Precision among the m highest-scored pairs (m = number of positives
    by default), ties broken by input order."""
    positive = np.asarray(positive, dtype=bool)
    m = int(positive.sum()) if m is None else int(m)
    order = np.argsort(-np.asarray(scores, dtype=np.float64), kind="stable")[:m]
    return float(np.count_nonzero(positive[order])) / m
