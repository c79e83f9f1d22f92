"""Writes tests/golden/ingest_*.json: dynamic-graph rows produced by the
reference's own ingest (Graph_dynamic.h, compiled by `make -C oracle ref` into
oracle/_ref/ref_graph).  Run in a container that holds the reference tree:

    python tests/golden/make_ingest_golden.py

Each fixture holds the snapshot files' text and, after every snapshot, each
vertex's degree, former degree and row, per direction.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_graph")


def run_ref(directed, n, texts):
    with tempfile.TemporaryDirectory() as td:
        paths = []
        for i, t in enumerate(texts):
            p = os.path.join(td, f"s{i}.txt")
            with open(p, "w") as f:
                f.write(t)
            paths.append(p)
        out = subprocess.run([REF_BIN, "d" if directed else "u", str(n)] + paths,
                             check=True, capture_output=True, text=True).stdout
    snaps = []
    for line in out.splitlines():
        tok = line.split()
        if tok[0] == "S":
            snaps.append({"m": int(tok[2]), "rows": [[None] * n for _ in range(2 if directed else 1)]})
        elif tok[0] == "V":
            d, v, deg, former = int(tok[1]), int(tok[2]), int(tok[3]), int(tok[4])
            row = [int(x) for x in tok[6:]]
            snaps[-1]["rows"][d][v] = [deg, former, row]
    return snaps


def edges_text(pairs, trailing_newline=True):
    s = "\n".join(f"{a} {b}" for a, b in pairs)
    return s + ("\n" if trailing_newline else "")


def fixtures():
    fx = []
    # hand-made: duplicates, a self-loop, an isolated vertex, both file endings
    t0 = edges_text([(0, 1), (1, 2), (2, 0), (3, 1), (1, 0)], True)
    t1 = edges_text([(0, 2), (4, 1), (2, 2)], False)
    t2 = edges_text([(4, 0), (0, 4), (1, 3)], True)
    for directed in (False, True):
        fx.append(("small_" + ("d" if directed else "u"), directed, 6, [t0, t1, t2]))
    # random multigraph snapshots
    rng = np.random.Generator(np.random.PCG64(11))
    for directed in (False, True):
        n = 40
        texts = []
        for k in range(4):
            e = rng.integers(0, n, size=(60 if k == 0 else 25, 2))
            texts.append(edges_text(e.tolist(), trailing_newline=(k % 2 == 0)))
        fx.append(("rand_" + ("d" if directed else "u"), directed, n, texts))
    return fx


def main():
    if not os.path.exists(REF_BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    for name, directed, n, texts in fixtures():
        snaps = run_ref(directed, n, texts)
        with open(os.path.join(HERE, f"ingest_{name}.json"), "w") as f:
            json.dump({"directed": directed, "n": n, "files": texts, "snapshots": snaps}, f)
        print("wrote", name, [s["m"] for s in snaps])


if __name__ == "__main__":
    sys.exit(main())
