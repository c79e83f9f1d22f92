"""Seeded synthetic workloads for the update->embedding path.

The reference runs on real snapshot files prepared by its own tools (subset
sampling Random_Sample_Points_LP.cpp:80-140, batch splitting
split_edges_NC.cpp:57-118).  These generators produce graphs of the same kind
in memory, deterministically from a seed:

* R-MAT (a, b, c, d) with optional node-id modulo, self-loops dropped,
  duplicates dropped (first occurrence kept), undirected edges canonicalised
  to (min, max);
* directed Chung-Lu with power-law out/in weights;
* a seeded shuffle cut into an initial snapshot and insertion batches
  (equal, lognormal, or a Poisson-arrival stream);
* the subset: distinct uniform draws among the edge sources of the initial
  snapshot, as Random_Sample_Points_LP.cpp:80-140 does.

All ids are int32; edge lists are (src, dst) pairs of int32 arrays.
"""
from __future__ import annotations

import dataclasses
import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _dedup_first(keys: np.ndarray) -> np.ndarray:
    """Indices of the first occurrence of every distinct key, in order."""
    _, first = np.unique(keys, return_index=True)
    first.sort()
    return first


def rmat_edges(scale: int, m: int, seed: int, abcd=(0.57, 0.19, 0.19, 0.05),
               directed: bool = False, n_mod: int | None = None,
               chunk: int = 1 << 24):
    """Exactly `m` distinct R-MAT edges over n = 2**scale (or n_mod) nodes."""
    a, b, c, _ = abcd
    n = n_mod if n_mod else (1 << scale)
    rng = _rng(seed)
    keys = np.empty(0, dtype=np.int64)
    while keys.size < m:
        need = m - keys.size
        batch = min(chunk, int(need * 1.25) + 4096)
        u = np.zeros(batch, dtype=np.int64)
        v = np.zeros(batch, dtype=np.int64)
        for _ in range(scale):
            r = rng.random(batch)
            bu = r >= (a + b)
            bv = ((r >= a) & (r < a + b)) | (r >= a + b + c)
            u = (u << 1) | bu
            v = (v << 1) | bv
        if n_mod:
            u %= n_mod
            v %= n_mod
        keep = u != v
        u, v = u[keep], v[keep]
        if not directed:
            u, v = np.minimum(u, v), np.maximum(u, v)
        k = np.concatenate([keys, u * n + v])
        keys = k[_dedup_first(k)]
    keys = keys[:m]
    return (keys // n).astype(np.int32), (keys % n).astype(np.int32), n


def chung_lu_edges(n: int, m: int, seed: int, out_exp: float, in_exp: float,
                   chunk: int = 1 << 24):
    """Exactly `m` distinct directed Chung-Lu edges (no self-loops)."""
    rng = _rng(seed)
    i = np.arange(n, dtype=np.float64)
    w_out = (i + 1.0) ** (-1.0 / (out_exp - 1.0))
    perm = rng.permutation(n)
    w_in = (perm + 1.0) ** (-1.0 / (in_exp - 1.0))
    c_out = np.cumsum(w_out); c_out /= c_out[-1]
    c_in = np.cumsum(w_in); c_in /= c_in[-1]
    keys = np.empty(0, dtype=np.int64)
    if m > 4 * chunk:
        # Large graphs: draws in one go, looked up as sorted queries (random
        # lookups in the CDF miss cache on every draw), the u sample randomly
        # matched to the v sample (still independent pairs); distinct keys,
        # then a random m of them.
        while keys.size < m:
            batch = int((m - keys.size) * 1.3) + 4096
            u = np.searchsorted(c_out, np.sort(rng.random(batch)), side="right").astype(np.int64)
            u = u[rng.permutation(batch)]
            v = np.searchsorted(c_in, np.sort(rng.random(batch)), side="right").astype(np.int64)
            u = np.minimum(u, n - 1); v = np.minimum(v, n - 1)
            keep = u != v
            keys = np.unique(np.concatenate([keys, u[keep] * n + v[keep]]))
        keys = keys[rng.permutation(keys.size)[:m]]
        return (keys // n).astype(np.int32), (keys % n).astype(np.int32)
    while keys.size < m:
        need = m - keys.size
        batch = min(chunk, int(need * 1.3) + 4096)
        u = np.searchsorted(c_out, rng.random(batch), side="right").astype(np.int64)
        v = np.searchsorted(c_in, rng.random(batch), side="right").astype(np.int64)
        u = np.minimum(u, n - 1); v = np.minimum(v, n - 1)
        keep = u != v
        k = np.concatenate([keys, u[keep] * n + v[keep]])
        keys = k[_dedup_first(k)]
    keys = keys[:m]
    return (keys // n).astype(np.int32), (keys % n).astype(np.int32)


def shuffle_split(src, dst, seed: int, init_frac: float, batch_sizes):
    """Seeded shuffle; the first init_frac of edges is the initial snapshot,
    then consecutive batches of the given sizes (truncated to what is left)."""
    rng = _rng(seed + 7919)
    perm = rng.permutation(src.size)
    s, d = src[perm], dst[perm]
    n0 = int(round(init_frac * src.size))
    snaps = [(s[:n0], d[:n0])]
    pos = n0
    for b in batch_sizes:
        if pos >= src.size:
            break
        b = int(min(b, src.size - pos))
        snaps.append((s[pos:pos + b], d[pos:pos + b]))
        pos += b
    return snaps


def lognormal_sizes(count: int, median: float, sigma: float, seed: int,
                    lo: int = 1, hi: int | None = None):
    rng = _rng(seed + 104729)
    x = np.exp(np.log(median) + sigma * rng.standard_normal(count))
    x = np.maximum(np.rint(x), lo)
    if hi is not None:
        x = np.minimum(x, hi)
    return x.astype(np.int64)


def sample_subset(init_src, nsub: int, seed: int):
    """Distinct uniform draws among the edge sources of the initial snapshot
    (Random_Sample_Points_LP.cpp:80-140)."""
    cand = np.unique(init_src)
    rng = _rng(seed + 15485863)
    if nsub > cand.size:
        raise ValueError("subset larger than the number of edge sources")
    return rng.choice(cand, size=nsub, replace=False).astype(np.int32)


@dataclasses.dataclass
class Workload:
    name: str
    n: int
    directed: bool
    snapshots: list            # [(src, dst)], snapshot 0 is the initial graph
    subset: np.ndarray
    alpha: float
    rmax: float
    nparts: int
    fanout: int
    rank: int
    q: int                     # frPCA passes (2 power iterations -> q = 6)
    fp64_state: bool           # PPR reserves/residues kept in FP64
    transform: str = "NC"      # "NC": log10(x/rmax) for x>rmax ; "LP": log10(1+x/rmax)
    delta: float = 0.65
    sparse_state: bool = False  # the n >= 1e7-node path's pair-list state (directed LP)
    merge_bf16x3: bool = False  # merge products as bf16x3 on MFMA (FP32 accumulation)


def workload(name: str, scale_down: int = 0, nsub: int | None = None,
             n_snap: int | None = None) -> Workload:
    """The named synthetic configurations.

    plain      undirected R-MAT, 131,072 nodes, 1M edges, 70% + 3x10%, |S|=256,
               alpha 0.2, rmax 1e-6, 16 blocks, rank 64, q=6, fan-out 2, FP64.
    batched    undirected R-MAT, 2,097,152 nodes, 16M edges, 60% + 12 lognormal
               batches (median 200k, sigma 1), |S|=3000, rmax 1e-7, 256 blocks,
               rank 128, q=6, fan-out 4, FP32 state.
    precision  directed Chung-Lu, 2,738,013 nodes, 13,960,811 edges, exponents
               out 2.1 / in 2.3, 80% + 8 equal batches, |S|=3000, rmax 1e-7,
               512 blocks (ragged fan-out-4 tree), rank 128.
    served     directed Chung-Lu, 6,216,199 nodes, 177,862,656 edges, exponents
               in 2.0 / out 2.3, 80% initial + 60 lognormal batches (median
               10k, sigma 1, clipped to [1k, 1M]), |S|=3000, rmax 1e-7, 1024
               blocks, rank 128, fan-out 4, LP transform.
    `scale_down` divides node and edge counts by 2**scale_down (tests only).
    """
    sd = 1 << scale_down
    if name == "plain":
        m = 1_000_000 // sd
        s, d, n = rmat_edges(17 - scale_down, m, seed=0)
        sizes = [m // 10] * 3
        snaps = shuffle_split(s, d, 0, 0.7, sizes)
        sub = sample_subset(snaps[0][0], nsub or 256, 0)
        return Workload(name, n, False, snaps, sub, 0.2, 1e-6, 16, 2, 64, 6, True)
    if name == "batched":
        m = 16_000_000 // sd
        s, d, n = rmat_edges(21 - scale_down, m, seed=1)
        k = n_snap if n_snap is not None else 12
        sizes = lognormal_sizes(k, 200_000 / sd, 1.0, 1)
        snaps = shuffle_split(s, d, 1, 0.6, sizes)
        sub = sample_subset(snaps[0][0], nsub or 3000, 1)
        return Workload(name, n, False, snaps, sub, 0.2, 1e-7, 256, 4, 128, 6, False)
    if name == "precision":
        n = 2_738_013 // sd
        m = 13_960_811 // sd
        s, d = chung_lu_edges(n, m, seed=2, out_exp=2.1, in_exp=2.3)
        rest = m - int(round(0.8 * m))
        sizes = [rest // 8 + (1 if i < rest % 8 else 0) for i in range(8)]
        snaps = shuffle_split(s, d, 2, 0.8, sizes)
        sub = sample_subset(snaps[0][0], nsub or 3000, 2)
        return Workload(name, n, True, snaps, sub, 0.2, 1e-7, 512, 4, 128, 6, False, merge_bf16x3=True)
    if name == "served":
        n = 6_216_199 // sd
        m = 177_862_656 // sd
        s, d = chung_lu_edges(n, m, seed=3, out_exp=2.3, in_exp=2.0)
        # 60 batches: the plain steps take the first few, the served stream the rest
        k = n_snap if n_snap is not None else 60
        sizes = lognormal_sizes(k, 10_000 / sd, 1.0, 3, max(1, 1_000 // sd), 1_000_000 // sd)
        snaps = shuffle_split(s, d, 3, 0.8, sizes)
        sub = sample_subset(snaps[0][0], nsub or 3000, 3)
        return Workload(name, n, True, snaps, sub, 0.2, 1e-7, 1024, 4, 128, 6, False, transform="LP")
    raise ValueError(f"unknown workload {name!r}")


def device_workload(capi, name: str, scale_down: int = 0, nsub: int | None = None):
    """The two largest configurations, drawn in HBM (their host generation
    takes minutes): returns (Workload whose snapshots are device slices of
    the generated edge arrays, the owned device arrays).

    served   directed Chung-Lu (tsvd_gen_chung_lu), 6,216,199 nodes,
             177,862,656 edges, out-exponent 2.3, in-exponent 2.0, seed 3;
             80% initial + 60 lognormal batches (median 10k, sigma 1, clipped
             to [1k, 1M]); |S| = 3000, rmax 1e-7, 1024 blocks, rank 128,
             fan-out 4, LP transform.
    atscale  directed R-MAT (0.57, 0.19, 0.19, 0.05) over 2^26 ids taken
             modulo 41,652,230 (tsvd_gen_rmat), 1,468,365,182 edges, seed 4;
             90% initial + 60 lognormal batches (median 100k, sigma 1);
             |S| = 3000, rmax 1e-7, 4096 blocks, rank 128, fan-out 4, LP.
    Sources are drawn among the edge sources of the first 16M initial edges
    (the edge order is random, so they sample the initial snapshot's sources
    uniformly)."""
    sd = 1 << scale_down
    if name == "served":
        n, m = 6_216_199 // sd, 177_862_656 // sd
        src, dst = capi.gen_chung_lu(n, m, 2.3, 2.0, seed=3)
        init, sizes, seed, nparts = 0.8, lognormal_sizes(60, 10_000 / sd, 1.0, 3, max(1, 1_000 // sd),
                                                          1_000_000 // sd), 3, 1024
    elif name == "atscale":
        n, m = 41_652_230 // sd, 1_468_365_182 // sd
        src, dst = capi.gen_rmat(int(np.ceil(np.log2(n))), n, m, 0.57, 0.19, 0.19, seed=4)
        init, sizes, seed, nparts = 0.9, lognormal_sizes(60, 100_000 / sd, 1.0, 4), 4, 4096
    else:
        raise ValueError(f"no device workload {name!r}")
    n0 = int(round(init * m))
    snaps = [(capi.DeviceSlice(src, 0, n0), capi.DeviceSlice(dst, 0, n0))]
    pos = n0
    for b in sizes:
        b = int(min(b, m - pos))
        if b <= 0:
            break
        snaps.append((capi.DeviceSlice(src, pos, b), capi.DeviceSlice(dst, pos, b)))
        pos += b
    head = capi.DeviceSlice(src, 0, min(n0, 1 << 24)).numpy()
    sub = sample_subset(head, nsub or 3000, seed)
    w = Workload(name, n, True, snaps, sub, 0.2, 1e-7, max(16, nparts // sd), 4, 128, 6, False, transform="LP")
    return w, (src, dst)


def request_stream(w: Workload, count: int, rate: float, seed: int, median: float = 10_000, sigma: float = 1.0,
                   lo: int = 1_000, hi: int = 1_000_000):
    """A served-mode stream over a workload's graph: after its initial
    snapshot, the remaining edges (in their shuffled order) arrive as `count`
    insertion batches with lognormal sizes (median 10k, sigma 1, clipped to
    [1k, 1M], the served configuration's stream) at Poisson times of the given
    rate (requests/s).  Returns ([(src, dst)], arrival offsets in seconds)."""
    s = np.concatenate([x[0] for x in w.snapshots[1:]])
    d = np.concatenate([x[1] for x in w.snapshots[1:]])
    sizes = lognormal_sizes(count, median, sigma, seed, lo, hi)
    out, pos = [], 0
    for b in sizes:
        if pos >= s.size:
            break
        b = int(min(b, s.size - pos))
        out.append((s[pos:pos + b], d[pos:pos + b]))
        pos += b
    rng = _rng(seed + 7)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(out)))
    arrivals -= arrivals[0]
    return out, arrivals
