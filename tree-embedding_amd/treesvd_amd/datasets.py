"""Dataset tools of the reference, restated on in-memory edge lists.

The reference prepares its inputs with three small programs whose outputs
the drivers then read (config.txt, subset file, snapshot edge files):

* subset sampling (Random_Sample_Points_LP.cpp:80-160): distinct uniform
  draws among the edge sources of the first snapshot file, written sorted,
  one id per line;
* link-prediction hold-out (gen_subset_LP_data_u.cpp:61-146 with
  UGraph::SubsetRandomSplitGraphWithTrainOutput / SubsetNegativeSamples,
  Graph_dynamic.h:665-723; the directed program gen_subset_LP_data_d.cpp
  with Graph's versions, :246-309): every snapshot is read into the graph,
  then a partial Fisher-Yates shuffle moves `percent` of its edges to the
  front; of those, the ones with an end in the subset become positive test
  pairs (subset end first), the rest of the snapshot is its training file;
  as many negative pairs are drawn as there are positives: a uniform subset
  vertex and a uniform vertex that is not its neighbour in the whole graph
  (out-neighbour when directed), with repeats allowed;
* batch splitting (split_edges_NC.cpp:57-118): the edges of consecutive
  snapshot files are cut into files of exactly 10,000 edges, at most 100 of
  them; a trailing remainder is dropped.

The reference draws with rand() seeded by time(0), so only the rules, not
the draws, can be matched; here every draw comes from a seeded PCG64.
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def sample_subset(first_src, count: int, seed: int = 0) -> np.ndarray:
    """Random_Sample_Points_LP.cpp:100-141: `count` distinct vertices drawn
    uniformly among the distinct edge sources of the first snapshot, sorted
    ascending as written to the subset file."""
    cand = np.unique(np.asarray(first_src))
    if count > cand.size:
        raise ValueError(f"subset of {count} from {cand.size} distinct sources")
    rng = _rng(seed)
    picked = np.zeros(cand.size, dtype=bool)
    out = []
    while len(out) < count:          # rejection of repeats, as check_if_selected_vec does
        i = int(rng.integers(cand.size))
        if not picked[i]:
            picked[i] = True
            out.append(int(cand[i]))
    return np.sort(np.asarray(out, dtype=np.int32))


def partial_shuffle(src, dst, percent: float, rng: np.random.Generator):
    """SubsetRandomSplitGraphWithTrainOutput (Graph_dynamic.h:665-686): for
    i < m2 = int(percent * m), swap edge i with edge i + r, r uniform in
    [0, m - i).  Returns the permuted copies and m2."""
    s = np.array(src, dtype=np.int32, copy=True)
    d = np.array(dst, dtype=np.int32, copy=True)
    m = s.size
    m2 = int(percent * m)
    for i in range(m2):
        r = int(rng.integers(m - i))
        j = i + r
        s[i], s[j] = s[j], s[i]
        d[i], d[j] = d[j], d[i]
    return s, d, m2


def lp_holdout(snapshots, subset, n: int, percent: float, directed: bool = False, seed: int = 0):
    """gen_subset_LP_data_{u,d}.cpp: returns (training snapshots, positive
    test pairs, negative test pairs) as lists of (src, dst) int32 arrays and
    two (count, 2) int32 arrays.  `snapshots` are all of the dataset's
    snapshot edge lists in order (the config's files after the subset file)."""
    rng = _rng(seed)
    in_s = np.zeros(n, dtype=bool)
    in_s[np.asarray(subset)] = True
    train, pos = [], []
    for src, dst in snapshots:
        s, d, m2 = partial_shuffle(src, dst, percent, rng)
        hs, hd = s[:m2], d[:m2]
        first = in_s[hs]                          # from in the subset: (from, to)
        second = ~first & in_s[hd]                # else to in the subset: (to, from)
        pos.append(np.stack([np.where(first, hs, hd)[first | second], np.where(first, hd, hs)[first | second]], 1))
        train.append((s[m2:].copy(), d[m2:].copy()))
    pos = np.concatenate(pos).astype(np.int32) if pos else np.zeros((0, 2), np.int32)
    # SubsetNegativeSamples: neighbours in the whole graph (every snapshot, test edges included)
    allsrc = np.concatenate([np.asarray(s) for s, _ in snapshots]).astype(np.int64)
    alldst = np.concatenate([np.asarray(d) for _, d in snapshots]).astype(np.int64)
    keys = allsrc * n + alldst
    if not directed:
        keys = np.concatenate([keys, alldst * n + allsrc])
    keys = np.unique(keys)
    sub = np.asarray(subset, dtype=np.int64)
    neg = []
    while len(neg) < pos.shape[0]:
        i = int(sub[int(rng.integers(sub.size))])
        j = int(rng.integers(n))
        k = i * n + j
        at = np.searchsorted(keys, k)
        if at < keys.size and keys[at] == k:
            continue
        neg.append((i, j))
    neg = np.asarray(neg, dtype=np.int32).reshape(-1, 2)
    return train, pos, neg


def split_batches(snapshots, group: int = 10_000, max_batches: int = 100):
    """split_edges_NC.cpp:57-118: the edges of the given snapshots, in order,
    cut into consecutive batches of exactly `group` edges, at most
    `max_batches` of them; edges left over after the last full batch are
    dropped."""
    s = np.concatenate([np.asarray(x, dtype=np.int32) for x, _ in snapshots]) if snapshots else np.zeros(0, np.int32)
    d = np.concatenate([np.asarray(y, dtype=np.int32) for _, y in snapshots]) if snapshots else np.zeros(0, np.int32)
    nb = min(max_batches, s.size // group)
    return [(s[i * group:(i + 1) * group].copy(), d[i * group:(i + 1) * group].copy()) for i in range(nb)]


def write_edges(path: str, src, dst):
    """A snapshot file: "from to" per line, newline-terminated (every tool of
    the reference writes `<< endl`, which is why its reader duplicates the
    last edge, Graph_dynamic.h:127-130)."""
    with open(path, "w") as f:
        for a, b in zip(np.asarray(src).tolist(), np.asarray(dst).tolist()):
            f.write(f"{a} {b}\n")


def write_pairs(path: str, pairs):
    with open(path, "w") as f:
        for a, b in np.asarray(pairs).reshape(-1, 2).tolist():
            f.write(f"{a} {b}\n")


def write_subset(path: str, subset):
    """Random_Sample_Points_LP.cpp:147-156: one id per line."""
    with open(path, "w") as f:
        for v in np.asarray(subset).tolist():
            f.write(f"{v}\n")


def write_config(path: str, subset_path: str, snapshot_paths):
    """config.txt (gen_subset_LP_data_u.cpp:131-141): the subset file, then
    the snapshot files in order, one path per line."""
    with open(path, "w") as f:
        f.write(subset_path + "\n")
        for p in snapshot_paths:
            f.write(p + "\n")
