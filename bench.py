#!/usr/bin/env python3
"""Update->embedding throughput of the Tree-SVD path on one MI355X.

A step is one snapshot of the plain per-snapshot method (tsvd_pipeline_snapshot):
an edge-insertion batch, already resident in HBM, goes through ingest ->
dynamic PPR update -> push -> block transform + lazy dirty test -> blockwise
frPCA of the dirty blocks -> dirty-path merges -> root embedding.  The initial
snapshot (static push over the first 60% of the edges) is setup and is timed
separately.  When warm-up + steps exceed the workload's insertion batches the
batches replay from a rebuilt initial state (outside the clock).  Then the
served form is compared with the plain method on request streams (the rest
of the graph arriving as lognormal 10k-edge batches at Poisson times): each
replays a stream on its own clock against a server, plain admitting one
request per snapshot, served admitting every queued request into the next
snapshot; at a rate above both capacities (capacities compared), at half the
plain method's capacity (steady-state p99 compared), and at saturation with
the two-sided threshold.  Output: one JSON line (snapshots/s, p50/p99 ms per
update, served vs plain, the dominant kernel's roofline, and the CPU
restatement timed on this host).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batched]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tree-embedding_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_MFMA_TFLOPS = 157.3       # MI355X dense FP32 matrix peak
FP64_MFMA_TFLOPS = 78.6
WORKLOAD_TAG = "batched"       # workload the committed counter summary was taken on


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pipeline_for(capi, w, seed=7):
    return capi.Pipeline(w.n, w.directed, w.subset, w.alpha, w.rmax, w.nparts, w.fanout, w.rank, w.q,
                         precision=capi.TSVD_FP64 if w.fp64_state else capi.TSVD_FP32,
                         transform=capi.TSVD_TRANSFORM_NC if w.transform == "NC" else capi.TSVD_TRANSFORM_LP,
                         delta=w.delta, dyn_start=2, lazy_start=2, slack=0.1, seed=seed,
                         sparse_state=w.sparse_state, merge_bf16x3=w.merge_bf16x3)


def roofline_of(prof, precision_fp64):
    """Dominant kernel of this library by total time (library solver calls,
    which carry no algorithmic count, are skipped); HBM-bound kernels report
    GB/s against 8 TB/s, the MFMA products TFLOP/s against the dense peak."""
    own = {k: v for k, v in prof.items() if v[2] > 0 or v[3] > 0}
    name, (launches, ms, nbytes, flops) = max(own.items(), key=lambda kv: kv[1][1])
    if name.startswith("gemm"):
        peak = FP64_MFMA_TFLOPS if precision_fp64 else FP32_MFMA_TFLOPS
        achieved = flops / (ms * 1e-3) / 1e12
        return {"kernel": name, "bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": None,
                "launches": launches, "avg_launch_ms": round(ms / launches, 4),
                "algorithmic_per_launch": flops / launches}
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "launches": launches, "avg_launch_ms": round(ms / launches, 4),
            "algorithmic_per_launch": nbytes / launches}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (two --pmc passes, FETCH_SIZE and WRITE_SIZE, calibrated on a copy
    of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"push_round_{WORKLOAD_TAG}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(w, gpu_ndirty, gpu_nmerged, n_steps_hint, budget_s=25.0):
    """The oracle (the CPU restatement of the reference path: C FIFO push on
    all host threads, numpy FP64 frPCA and merges) on a bounded sample of the
    same workload: the push/update of a sample of sources over the initial
    snapshot and one insertion batch, scaled by |S|/sample (sources are
    independent), plus one full-size frPCA and one merge, scaled by the
    device's dirty-block and merged-node counts per step."""
    from oracle import oracle as orc
    # the box's CPU share is 16 threads, whatever os.cpu_count() reports
    nthreads = min(os.cpu_count() or 1, 16)
    nsamp = min(w.subset.size, max(8, 2 * nthreads))
    sub = w.subset[:nsamp]
    g = orc.Graph(w.n, w.directed)
    s0, d0 = w.snapshots[0]
    g.append(s0, d0)
    ppr = orc.PPR(w.n, sub, w.directed, w.alpha, w.rmax, fp32=not w.fp64_state)
    ppr.init_static()
    ppr.push(g, nthreads)
    s1, d1 = w.snapshots[1]
    t0 = time.perf_counter()
    g.append(s1, d1)
    ppr.refresh_dynamic(g, nthreads)
    ppr.push(g, nthreads)
    t_push = (time.perf_counter() - t0) * (w.subset.size / nsamp)
    starts, widths = orc.block_map(w.n, w.nparts)
    rng = np.random.Generator(np.random.PCG64(1))
    A = np.abs(rng.standard_normal((w.subset.size, int(widths[0])))) * 3.0
    t1 = time.perf_counter()
    orc.frpca(A, w.rank, w.q, omega=orc.sketch(3, A.shape[1], w.rank + 5))
    t_frpca = time.perf_counter() - t1
    C = rng.standard_normal((w.subset.size, w.fanout * w.rank))
    t2 = time.perf_counter()
    orc.merge([C], w.rank)
    t_merge = time.perf_counter() - t2
    per_snap = t_push + gpu_ndirty * t_frpca + gpu_nmerged * t_merge
    return {"value": round(1.0 / per_snap, 6), "unit": "snapshots/s", "cores": nthreads, "kind": "port",
            "seconds_per_snapshot": round(per_snap, 3),
            "sample": (f"push+update of {nsamp}/{w.subset.size} sources over one insertion batch "
                       f"({s1.size} edges) scaled x{w.subset.size / nsamp:.1f}; one {A.shape[0]}x{A.shape[1]} "
                       f"FP64 frPCA (k={w.rank}, q={w.q}) x {gpu_ndirty:.1f} dirty blocks; one "
                       f"{C.shape[0]}x{C.shape[1]} merge SVD x {gpu_nmerged:.1f} merges"),
            "phase_s": {"push_update": round(t_push, 3), "frpca_one_block": round(t_frpca, 3),
                        "merge_one": round(t_merge, 3)}}


class DenseOracle:
    """Exact top-k eigenpairs of G = A A^T for the |S| x n transformed matrix A
    in FP64.  Small cases form G on the device and call numpy's eigh; large
    ones run block subspace iteration on G through gram_apply (A (A^T Y), no
    Gram) until every top-k Ritz pair's residual ||G u - theta u|| is below
    `tol` * theta: a symmetric matrix then has an eigenvalue within that
    residual of theta, so the reported bound certifies the sigma values."""

    def __init__(self, pipe, k, a_norms, gram_limit=4e13, tol=1e-9, max_iter=120, seed=11):
        self.pipe, self.k = pipe, k
        ns = pipe.nsrc
        t = time.perf_counter()
        if float(ns) * ns * pipe.n <= gram_limit:
            self.kind = "FP64 Gram A A^T over all columns (device MFMA) + numpy eigh"
            self.G = pipe.dense_gram()
            lam, V = np.linalg.eigh(self.G)
            self.lam, self.U = lam[::-1][:k].copy(), V[:, ::-1][:, :k].copy()
            self.tr = float(np.trace(self.G))
            self.bound, self.iters = 0.0, 0
        else:
            self.G = None
            self.tr = float(np.sum(np.asarray(a_norms, dtype=np.float64) ** 2))
            l = min(ns, 2 * k + 64)
            rng = np.random.Generator(np.random.PCG64(seed))
            Y, _ = np.linalg.qr(rng.standard_normal((ns, l)))
            for it in range(1, max_iter + 1):
                Z = pipe.gram_apply(Y)
                H = Y.T @ Z
                th, W = np.linalg.eigh(0.5 * (H + H.T))
                th, W = th[::-1], W[:, ::-1]
                R = Z @ W[:, :k] - (Y @ W[:, :k]) * th[None, :k]
                rel = np.linalg.norm(R, axis=0) / np.maximum(th[:k], 1e-300)
                self.lam, self.U = th[:k].copy(), Y @ W[:, :k]
                self.bound, self.iters = float(rel.max()), it
                log(f"oracle subspace iteration {it}: l={l} max residual/theta over top {k} = {self.bound:.2e} "
                    f"({time.perf_counter() - t:.1f}s)")
                if self.bound < tol:
                    break
                Y, _ = np.linalg.qr(Z)
            self.kind = (f"FP64 block subspace iteration on A A^T (device MFMA products, l={l}, {self.iters} "
                         f"iterations, max top-{k} residual/theta {self.bound:.1e})")
        self.sigma = np.sqrt(np.clip(self.lam, 0, None))
        self.seconds = time.perf_counter() - t

    def apply(self, U):
        return self.G @ U if self.G is not None else self.pipe.gram_apply(U)

    def errors(self, sig, E, mode_nc):
        """sigma relative errors, subspace error 1 - ||U_o^T U||_F^2 / k and
        reconstruction ||A - U U^T A||_F^2 / ||A||_F^2 = 1 - tr(U^T G U) / tr(G)
        of an embedding (NC: U sqrt(S), LP: U) against the oracle."""
        k = self.k
        U = E / np.sqrt(np.where(sig > 0, sig, 1.0))[None, :] if mode_nc else E
        U, _ = np.linalg.qr(U)
        rel = np.abs(sig - self.sigma) / self.sigma
        sub = 1.0 - float(np.linalg.norm(self.U.T @ U) ** 2) / k
        rec = 1.0 - float(np.sum(U * self.apply(U))) / self.tr
        best = 1.0 - float(self.lam.sum()) / self.tr
        return {"sigma_rel_err_max": float(rel.max()), "sigma_rel_err_top16_max": float(rel[:16].max()),
                "sigma_rel_err_median": float(np.median(rel)), "subspace_err": sub,
                "reconstruction_err": rec, "reconstruction_err_optimal_rank_k": best}


def accuracy_report(pipe, w, seed=7):
    """Root singular values, subspace and reconstruction error of the device
    embedding against the exact FP64 SVD of the current |S| x n matrix, next
    to the reference algorithm's own error: the oracle's FP64 frPCA (frpca.c)
    and exact merge SVDs (Tree_u_NC.cpp:723-850) applied to the same stored
    blocks."""
    from oracle import oracle as orc
    k = w.rank
    nc = w.transform == "NC"
    a_norms = pipe.block_norms()[0]
    orac = DenseOracle(pipe, k, a_norms)
    log(f"oracle ready in {orac.seconds:.1f}s: {orac.kind}")
    dev = orac.errors(pipe.root_sigma(), pipe.embedding(), nc)
    t = time.perf_counter()
    starts, widths = orc.block_map(w.n, w.nparts)
    leaves = []
    for j in range(w.nparts):
        Aj = pipe.block_matrix(j, int(widths[j]))
        # the block's stored columns (active at its last rebuild): the sketch
        # has one row per stored column, as on the device
        cols = pipe.block_columns(j) - starts[j]
        Ac = Aj[:, cols]
        del Aj
        if not Ac.any():
            leaves.append(np.zeros((w.subset.size, k)))
        else:
            U, S, _ = orc.frpca(Ac, k, w.q, omega=orc.sketch(orc.leaf_seed(seed, j), cols.size, k + 5))
            leaves.append(U * S)
        if (j + 1) % 32 == 0 or j + 1 == w.nparts:
            log(f"reference algorithm: {j + 1}/{w.nparts} leaf frPCAs ({time.perf_counter() - t:.1f}s)")
    levels = orc.tree_levels(w.nparts, w.fanout)
    cur = leaves
    for lvl in range(1, len(levels)):
        groups = [cur[i * w.fanout:min((i + 1) * w.fanout, len(cur))] for i in range(levels[lvl])]
        if lvl == len(levels) - 1:
            _, Ur, Sr = orc.merge(groups[0], k)
        else:
            cur = [orc.merge(ch, k)[0] for ch in groups]
        log(f"reference algorithm: merge level {lvl} ({levels[lvl]} nodes)")
    if len(levels) == 1:
        _, Ur, Sr = orc.merge(leaves, k)
    ref = orac.errors(Sr, Ur * np.sqrt(Sr)[None, :] if nc else Ur, nc)
    t_ref = time.perf_counter() - t
    gap = np.abs(pipe.root_sigma() - Sr) / Sr
    return {"device": dev, "reference_algorithm_fp64": ref,
            "device_vs_reference_sigma_rel_max": float(gap.max()),
            "device_vs_reference_sigma_rel_top16_max": float(gap[:16].max()),
            "dense_oracle": orac.kind, "oracle_sigma_rel_bound": 0.5 * orac.bound,
            "seconds": {"oracle": round(orac.seconds, 1), "reference_algorithm": round(t_ref, 1)}}


def served_run(capi, w, reqs, arrivals, cap, label, seed=7, overlap=True, push_setting=None, two_sided=False,
               window=None):
    """Replays a request stream against a server on its own clock: a feeder
    thread submits request i at arrivals[i] seconds, this thread collects the
    results in order.  cap = 1 is the plain per-snapshot method (steps one at
    a time), cap = 0 the served form (every queued request admitted into the
    next step, the back part of each step overlapping the next step's push).
    Latency percentiles cover the requests that arrive inside `window`
    (seconds from the first arrival; default: all of them).  Returns the
    metrics and the root singular values of the final state."""
    import threading
    srv = capi.Server(w.n, w.directed, w.subset, w.alpha, w.rmax, w.nparts, w.fanout, w.rank, w.q,
                      precision=capi.TSVD_FP64 if w.fp64_state else capi.TSVD_FP32,
                      transform=capi.TSVD_TRANSFORM_NC if w.transform == "NC" else capi.TSVD_TRANSFORM_LP,
                      delta=w.delta, dyn_start=2, lazy_start=2, slack=0.1, seed=seed, max_step_requests=cap,
                      sparse_state=w.sparse_state, overlap=overlap and cap != 1, two_sided=two_sided,
                      merge_bf16x3=w.merge_bf16x3)
    t = time.perf_counter()
    if push_setting is not None:          # the plain run's autotuned push setting, for both replays
        srv.pipeline.ppr.push_setting = push_setting
    s0, d0 = w.snapshots[0]
    if isinstance(s0, capi.DeviceSlice):
        srv.pipeline.snapshot(s0, d0)     # device-resident initial graph; the worker is idle until a submit
    else:
        srv.wait(srv.submit(s0, d0), embedding=False)
    log(f"{label}: initial snapshot {time.perf_counter() - t:.1f}s")
    tickets = [None] * len(reqs)
    ready = [threading.Event() for _ in reqs]
    pr0 = capi.profile_read()
    start = time.perf_counter()

    def feeder():
        for i, (s, d) in enumerate(reqs):
            delay = start + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(s, d)
            ready[i].set()

    th = threading.Thread(target=feeder)
    th.start()
    res = []
    for i in range(len(reqs)):
        ready[i].wait()
        _, r = srv.wait(tickets[i], embedding=False)
        res.append(r)
        if r["last_ticket"] == r["ticket"]:
            ph = r["step"]
            log(f"{label}: step {r['snapshot']} done: {r['step_requests']} requests, {r['step_edges']} edges, "
                f"{r['done_ms'] - r['start_ms']:.0f} ms, latency of request {i}: {r['latency_ms']:.0f} ms "
                f"(ingest {ph['ms_ingest']:.0f}, push {ph['ms_push']:.0f}, blocks {ph['ms_blocks']:.0f}, "
                f"frpca {ph['ms_frpca']:.0f}, merge {ph['ms_merge']:.0f} ms)")
    th.join()
    pr1 = capi.profile_read()

    def delta(name):
        a, b = pr0.get(name, (0, 0.0)), pr1.get(name, (0, 0.0))
        return f"{b[0] - a[0]} ({b[1] - a[1]:.0f} ms)"
    log(f"{label}: CSR rebuilds {delta('graph_rebuild')}, hipMalloc {delta('hip_malloc')}, "
        f"hipFree {delta('hip_free')} over the stream")
    sigma = srv.pipeline.root_sigma()
    srv.close()
    inside = [i for i in range(len(res)) if window is None or window[0] <= arrivals[i] <= window[1]]
    lat = np.array([res[i]["latency_ms"] for i in inside])
    span = (max(r["done_ms"] for r in res) - min(r["submit_ms"] for r in res)) / 1e3
    steps = sorted({r["snapshot"] for r in res})
    return {"requests_per_s": round(len(res) / span, 4), "device_snapshots_per_s": round(len(steps) / span, 4),
            "p50_ms": round(float(np.percentile(lat, 50)), 1), "p99_ms": round(float(np.percentile(lat, 99)), 1),
            "mean_ms": round(float(lat.mean()), 1), "latency_requests": len(inside),
            "steps": len(steps), "requests_per_step": round(len(res) / len(steps), 2),
            "makespan_s": round(span, 2)}, sigma


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="batched", choices=["plain", "batched", "precision", "served", "atscale"])
    ap.add_argument("--transform", choices=["NC", "LP"], help="override the workload's block transform")
    ap.add_argument("--shard", default=None, metavar="I/N",
                    help="run GPU I's share of an N-GPU source-sharded job: the I-th of N contiguous "
                         "slices of the subset")
    ap.add_argument("--sparse-state", action="store_true",
                    help="the n >= 1e7-node path: pair-list state, pushed a batch of source groups at a time "
                         "(directed workloads, LP transform)")
    ap.add_argument("--scale-down", type=int, default=0, help="divide nodes/edges by 2**k (smoke runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--accuracy", action="store_true",
                    help="append errors against a dense FP64 SVD and the reference algorithm's own error")
    ap.add_argument("--served-requests", type=int, default=48,
                    help="requests in the served-vs-plain stream (0: skip that comparison)")
    ap.add_argument("--served-rate", type=float, default=24.0,
                    help="lowest Poisson arrival rate of the saturation streams (1/s); the rate used is at least "
                         "4x the plain method's measured snapshots/s times the ratio of the workload's median "
                         "batch to the median request, above what either method sustains, so that each runs "
                         "flat out")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the push autotuner on the first warm-up snapshot and the MFMA tile-width "
                         "autotuner over the initial and warm-up snapshots")
    ap.add_argument("--served-two-sided", action=argparse.BooleanOptionalAction, default=True,
                    help="also compare at saturation with the two-sided push threshold")
    ap.add_argument("--served-no-overlap", action="store_true",
                    help="run the served form's steps one at a time (no back/front overlap)")
    ap.add_argument("--push-setting", default=None, metavar="HUB,RIDE,K",
                    help="fixed push setting (hub chunk, rider threshold, items per wave) for every pipeline; "
                         "no push autotuning")
    ap.add_argument("--push-graphs", action="store_true",
                    help="replay the push rounds from captured hipGraphs (off by default: measured slower)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="first run one device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; --gpus must be 1 "
                          "(source sharding over several GPUs is designed, not launched; see --shard)"}))
        sys.exit(2)

    if args.push_graphs:
        os.environ["TSVD_PUSH_GRAPHS"] = "1"   # read when each PPR handle is created
    from treesvd_amd import capi, gen
    t = time.time()
    owned = None
    if args.workload in ("served", "atscale"):
        w, owned = gen.device_workload(capi, args.workload, args.scale_down)
        args.no_cpu_baseline = True   # the CPU leg would need the 1.8e8 / 1.47e9-edge graph on the host
    else:
        w = gen.workload(args.workload, scale_down=args.scale_down)
    if args.transform:
        w.transform = args.transform
    shard = None
    if args.shard:
        i, nsh = (int(x) for x in args.shard.split("/"))
        if not 0 <= i < nsh:
            raise SystemExit("--shard I/N needs 0 <= I < N")
        w.subset = np.array_split(w.subset, nsh)[i].copy()
        shard = {"index": i, "of": nsh, "sources": int(w.subset.size)}
        log(f"shard {i}/{nsh}: {w.subset.size} sources")
    if args.sparse_state:
        if not w.directed:
            raise SystemExit("--sparse-state is the directed LP path")
        w.transform, w.sparse_state = "LP", True
    log(f"workload {args.workload}: n={w.n} snapshots={len(w.snapshots)} |S|={w.subset.size} "
        f"gen {time.time() - t:.1f}s")
    if args.steps < 1 or args.warmup < 0:
        raise SystemExit("--steps must be >= 1 and --warmup >= 0")
    nbatch = len(w.snapshots) - 1
    if nbatch < 1:
        raise SystemExit("workload has no insertion batch")
    name, hbm_total, hbm_free = capi.device_info()
    log(f"device {name}, HBM {hbm_total / 2**30:.0f} GiB ({hbm_free / 2**30:.0f} free)")

    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        a = capi.DeviceArray(np.zeros(nb // 4, np.float32))
        b = capi.DeviceArray(np.ones(nb // 4, np.float32))
        capi.check(capi.lib().tsvd_memcpy(a.ptr, b.ptr, nb, 2))
        log(f"calibration copy of {args.calibrate_copy} MiB done")
        a.free(); b.free()
    def on_device(x):
        return x if isinstance(x, capi.DeviceSlice) else capi.DeviceArray(x.astype(np.int32))

    batches = [(on_device(s), on_device(d)) for s, d in w.snapshots]
    # kernel events cover every launch of the run (as a rocprofv3 trace of this
    # command does); the snapshot rate comes from the timed steps only
    capi.profile_reset()
    capi.profile_enable(True)
    if not args.no_tune:
        # product tile widths chosen on the initial snapshot (every block refactored,
        # every merge recomputed) and the warm-up steps, kept for the timed steps
        capi.gemm_tuning(True)
    pipe = pipeline_for(capi, w)
    fixed = None
    if args.push_setting:
        h, r, k = args.push_setting.split(",")
        fixed = (int(h), float(r), int(k))
        pipe.ppr.push_setting = fixed
    t = time.perf_counter()
    st0 = pipe.snapshot(*batches[0])
    init_ms = (time.perf_counter() - t) * 1e3
    log(f"initial snapshot {init_ms:.0f} ms (push {st0['ms_push']:.0f} ms, {st0['push_rounds']} rounds, "
        f"frpca {st0['ms_frpca']:.0f} ms, {st0['ndirty']} blocks)")
    if w.sparse_state:
        log("sparse state resident (no pair lists)" if st0["state_batches"] == 0 else
            f"pair-list state: {st0['state_entries']:.4e} entries ({st0['state_entries'] * 8 / 1e9:.2f} GB), "
            f"{st0['state_batches']} source batches")
    # The workload has `nbatch` insertion batches (12 for the batched graph).
    # When warm-up + timed steps need more, the stream is replayed: the
    # pipeline is rebuilt from the initial snapshot (outside the clock) and
    # the batches run again from the first one, so every timed step is one of
    # the configuration's own insertion snapshots applied to its own
    # predecessor state.
    cur = {"pos": 1, "replays": [], "push_setting": fixed}

    def next_batch():
        nonlocal pipe
        if cur["pos"] > nbatch:
            pipe.close()
            t0 = time.perf_counter()
            pipe = pipeline_for(capi, w)
            if cur["push_setting"] is not None:
                pipe.ppr.push_setting = cur["push_setting"]
            pipe.snapshot(*batches[0])
            cur["replays"].append(round((time.perf_counter() - t0) * 1e3, 1))
            log(f"replay {len(cur['replays'])}: rebuilt from the initial snapshot in {cur['replays'][-1]:.0f} ms")
            cur["pos"] = 1
        j = cur["pos"]
        cur["pos"] += 1
        return j

    warm = []
    tuning = None
    for i in range(args.warmup):
        j = next_batch()
        if i == 0 and not args.no_tune and fixed is None:
            # the push autotuner: this snapshot's push timed under each setting from the same state
            cands = [(h, 0.5, k) for h in (128, 256, 4096, 8192) for k in (1, 8)]
            st, chosen, ms = pipe.snapshot_tuned(*batches[j], cands)
            tuning = {"candidates": [{"hub_chunk": h, "ride_frac": r, "items_per_wave": k,
                                      "push_ms": round(float(t), 1)} for (h, r, k), t in zip(cands, ms)],
                      "chosen": {"hub_chunk": chosen[0], "ride_frac": chosen[1], "items_per_wave": chosen[2]}}
            cur["push_setting"] = tuple(chosen)
            log(f"push autotuner: chose hub_chunk {chosen[0]}, ride_frac {chosen[1]}, items/wave {chosen[2]} "
                f"({', '.join(f'{h}/{r}/{k}: {t:.0f}' for (h, r, k), t in zip(cands, ms))} ms)")
        else:
            st = pipe.snapshot(*batches[j])
        warm.append(st)
        log(f"warmup {i} (batch {j}): {st['ms_total']:.1f} ms dirty={st['ndirty']}")
    capi.gemm_tuning(False)
    gemm_tuning = capi.gemm_tuning_read()
    if gemm_tuning:
        log("tile autotuner: " + ", ".join(f"{k.split('|')[0]} N{k.split('|')[2][1:]} -> {v['tile_n']} "
                                           f"({v['ms_tile80']:.2f}/{v['ms_tile144']:.2f} ms)"
                                           for k, v in gemm_tuning.items()))

    # K timed steps: each is one synchronous snapshot call (the library
    # synchronises its streams before returning), timed on the host clock;
    # the push's kernel time is read around each step so that replays'
    # initial snapshots stay out of the push rate
    lat, stats, step_batch = [], [], []
    push_ms = 0.0
    push_kernels = ("ppr_push_round", "ppr_push_hub")
    for i in range(args.steps):
        j = next_batch()
        before = capi.profile_read()
        t = time.perf_counter()
        st = pipe.snapshot(*batches[j])
        lat.append((time.perf_counter() - t) * 1e3)
        after = capi.profile_read()
        push_ms += sum(after.get(k, (0, 0.0))[1] - before.get(k, (0, 0.0))[1] for k in push_kernels)
        stats.append(st)
        step_batch.append(j)
        log(f"step {i} (batch {j}): {lat[-1]:.1f} ms dirty={st['ndirty']} merged={st['nmerged']} "
            f"push {st['ms_push']:.1f} ({st['push_rounds']} rounds) blocks {st['ms_blocks']:.1f} "
            f"frpca {st['ms_frpca']:.1f} merge {st['ms_merge']:.1f} pops {st['push_pops']} relax {st['push_relax']} "
            f"items {st['push_items']} wave-relax {st['push_wave_relax']}")
    total_s = sum(lat) * 1e-3
    value = args.steps / total_s
    # the push's own ceiling: its work is 64-lane atomic adds, so its rate in
    # wave-level atomics (row visits of popped items, round + hub kernels) over
    # the timed steps is set against what a pure stream of such atomics
    # reaches on this device
    wave_atomics = sum(s["push_wave_relax"] for s in stats)
    line_atomics = sum(s["push_line_relax"] for s in stats)
    ceiling = capi.probe_atomic_rate()
    probe_noret = capi.probe_atomic_lanes(returning=False, lanes=64)
    probe_line = capi.probe_atomic_lanes(returning=True, lanes=32)     # one 128-byte line per instruction
    achieved_rate = wave_atomics / (push_ms * 1e-3) if push_ms > 0 else 0.0
    line_rate = line_atomics / (push_ms * 1e-3) if push_ms > 0 else 0.0
    log(f"push: {wave_atomics:.3e} wave-atomics in {push_ms:.0f} ms = {achieved_rate / 1e9:.2f} G/s; "
        f"probe ceiling {ceiling / 1e9:.2f} G/s (non-returning {probe_noret / 1e9:.2f}); "
        f"{line_rate / 1e9:.2f} G lines/s against {probe_line / 1e9:.2f} G lines/s for one-line atomics")
    graph_replays, graph_captures, graph_failures = pipe.ppr.graph_stats()
    ndirty = float(np.mean([s["ndirty"] for s in stats]))
    nmerged = float(np.mean([s["nmerged"] for s in stats]))
    phases = {k: round(float(np.mean([s[k] for s in stats])), 3)
              for k in ("ms_ingest", "ms_refresh", "ms_push", "ms_blocks", "ms_frpca", "ms_merge")}
    state = None
    if w.sparse_state:
        ent = stats[-1]["state_entries"]
        dense_gb = round(4 * w.n * (-(-w.subset.size // 64) * 64) * 4 / 1e9, 3)
        if stats[-1]["state_batches"] == 0:
            state = {"resident": True, "dense_GB": dense_gb,
                     "note": "the whole state fits in HBM: one resident batch, no pair lists"}
        else:
            state = {"batches": stats[-1]["state_batches"], "entries": int(ent), "GB": round(ent * 8 / 1e9, 3),
                     "dense_equivalent_GB": dense_gb,
                     "note": "positive (vertex, FP32 value) pairs of p and r, both directions, all sources"}
    out = {
        "metric": "update->embedding snapshots/s (plain GPU method), p50/p99 ms per update; served vs plain",
        "value": round(value, 4), "unit": "snapshots/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(total_s / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if w.fp64_state else "f32", "data": "synthetic",
        "config": {"workload": args.workload, "n": w.n, "directed": w.directed,
                   "edges": int(sum(s.size for s, _ in w.snapshots)), "sources": int(w.subset.size),
                   "alpha": w.alpha, "rmax": w.rmax, "nparts": w.nparts, "fanout": w.fanout, "rank": w.rank,
                   "q": w.q, "scale_down": args.scale_down, "method": "plain", "transform": w.transform,
                   "sparse_state": w.sparse_state, "shard": shard},
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "initial_snapshot_ms": round(init_ms, 1), "phase_ms": phases,
        "timed_batches": step_batch, "insertion_batches": nbatch,
        "replays": {"count": len(cur["replays"]), "rebuild_ms": cur["replays"],
                    "note": "when warm-up + steps exceed the workload's insertion batches, the pipeline is rebuilt "
                            "from the initial snapshot outside the clock and the batches replay from the first"},
        "dirty_blocks_per_step": ndirty, "merged_nodes_per_step": nmerged, "pair_list_state": state,
        "norm_items_per_step": float(np.mean([s["norm_items"] for s in stats])),
        "push_graphs": {"replays": graph_replays, "captures": graph_captures, "failed_captures": graph_failures,
                        "enabled": bool(args.push_graphs),
                        "note": "with --push-graphs the push rounds are replayed from captured hipGraphs, six "
                                "rounds per replay; counted over the last pipeline (rebuilt at each replay of the "
                                "batches)"},
        "push_tuning": tuning,
        "gemm_tile_tuning": gemm_tuning,
        "push_atomic_roofline": {"achieved": round(achieved_rate / 1e9, 3), "ceiling": round(ceiling / 1e9, 3),
                                 "unit": "G wave-atomics/s", "frac": round(achieved_rate / ceiling, 4),
                                 "ceiling_non_returning": round(probe_noret / 1e9, 3),
                                 "lines": {"achieved": round(line_rate / 1e9, 3),
                                           "ceiling": round(probe_line / 1e9, 3), "unit": "G lines/s",
                                           "frac": round(line_rate / probe_line, 4)},
                                 "note": "plain-method push over the timed steps (round + hub kernels); ceiling = "
                                         "64-lane FP32 atomics at hashed 256-byte rows (tsvd_probe_atomic_rate; "
                                         "returning and non-returning run at the same rate); lines = 128-byte "
                                         "lines the relaxation atomics touch against 32-lane (one-line) atomics, "
                                         "the unit the L2 atomic pipes are paced by"},
    }
    if args.accuracy:
        log("accuracy against the dense FP64 oracle ...")
        out["accuracy"] = accuracy_report(pipe, w)
    pipe.close()
    for b in batches:
        for x in b:
            x.free()
    if args.served_requests > 0:
        if owned is not None:   # the stream's batches on the host; the initial snapshot stays on the device
            import dataclasses
            w = dataclasses.replace(w, snapshots=[w.snapshots[0]] + [(s.numpy(), d.numpy())
                                                                     for s, d in w.snapshots[1:]])
        # request sizes of the workload's stream: median 100k at the 1.47e9-edge shape, 10k otherwise
        med = 100_000 if args.workload == "atscale" else 10_000
        ps = ((tuning["chosen"]["hub_chunk"], tuning["chosen"]["ride_frac"], tuning["chosen"]["items_per_wave"])
              if tuning else None)

        def compare(name, rate, count, two_sided, window_frac=None):
            """One stream replayed by the plain method and by the served form."""
            reqs, arr = gen.request_stream(w, count, rate, seed=11, median=med / (1 << args.scale_down))
            window = None
            if window_frac is not None:   # latencies of the arrivals after the first warm-up part
                window = (window_frac * arr[-1], arr[-1])
            log(f"served vs plain [{name}]: {len(reqs)} requests ({sum(s.size for s, _ in reqs)} edges, lognormal "
                f"median {med / (1 << args.scale_down):.0f}) at {rate:.2f}/s Poisson, "
                f"{'two-sided' if two_sided else 'reference'} threshold")
            plain, s_plain = served_run(capi, w, reqs, arr, 1, f"plain [{name}]", push_setting=ps,
                                        two_sided=two_sided, window=window)
            served, s_served = served_run(capi, w, reqs, arr, 0, f"served [{name}]",
                                          overlap=not args.served_no_overlap, push_setting=ps,
                                          two_sided=two_sided, window=window)
            gap = np.abs(s_served - s_plain) / np.maximum(s_plain, 1e-300)
            return {"rate_per_s": round(rate, 3), "requests": len(reqs), "edges": int(sum(s.size for s, _ in reqs)),
                    "threshold": "two-sided |r| > rmax" if two_sided else "reference r > rmax (ppr:1445)",
                    "latency_window_s": None if window is None else [round(window[0], 2), round(window[1], 2)],
                    "plain": plain, "served": served,
                    "throughput_ratio": round(served["requests_per_s"] / plain["requests_per_s"], 2),
                    "p99_ratio": round(served["p99_ms"] / plain["p99_ms"], 4),
                    "final_sigma_served_vs_plain": {"top16_max_rel": float(gap[:16].max()),
                                                    "all_max_rel": float(gap.max())}}

        rows = {}
        # capacity: an arrival rate above what either method sustains, so each
        # runs flat out and its requests/s is its capacity (the p99 of an
        # overloaded queue grows with the stream's length, so it is not a
        # steady-state figure)
        # above both capacities: 4x the plain method's snapshot rate, scaled by
        # how much smaller the requests are than the batches that rate was
        # measured on (a step's cost grows with its edges)
        batch_med = float(np.median([x.size for x, _ in w.snapshots[1:]])) if len(w.snapshots) > 1 else med
        sat_rate = max(args.served_rate, 4.0 * value * max(1.0, batch_med / (med / (1 << args.scale_down))))
        rows["saturation"] = compare("saturation", sat_rate, args.served_requests, False)
        plain_cap = rows["saturation"]["plain"]["requests_per_s"]
        # steady state: half the plain method's capacity, latencies of the
        # arrivals after the first quarter of the stream (both queues stable)
        steady_rate = 0.5 * plain_cap
        rows["steady"] = compare("steady", steady_rate, max(12, args.served_requests // 2), False, window_frac=0.25)
        if args.served_two_sided:
            rows["saturation_two_sided"] = compare("saturation, two-sided", sat_rate, args.served_requests,
                                                   True)
        out["served_vs_plain"] = {
            "stream": {"sizes": f"lognormal median {med // 1000}k, sigma 1, clip [1k, 1M]",
                       "graph": f"{args.workload} workload after its initial snapshot", "arrivals": "Poisson"},
            "served_overlap": not args.served_no_overlap, **rows,
            "note": ("requests_per_s = requests completed / makespan; device_snapshots_per_s = pipeline steps / "
                     "makespan; saturation compares capacities, steady compares p99 at a rate both sustain; "
                     "final_sigma_served_vs_plain compares the root singular values after the whole stream, "
                     "grouped one request per step (plain) or as the served form batched them")}
    # kernel events cover every launch of this run, as a rocprofv3 trace of the
    # same command does
    capi.profile_enable(False)
    prof = capi.profile_read()
    roof = roofline_of(prof, w.fp64_state)
    if args.workload == WORKLOAD_TAG and args.scale_down == 0 and roof["kernel"] == "ppr_push_round":
        roof["traffic"] = profiled_traffic("k_push_round")
        roof["traffic_source"] = f"profiles/push_round_{WORKLOAD_TAG}_hbm_traffic.md"
    out["roofline"] = roof
    host = ("hip_malloc", "hip_free", "graph_rebuild")   # host wall times, not kernels
    out["kernels"] = {k: {"launches": v[0], "ms": round(v[1], 3), "GB": round(v[2] / 1e9, 3),
                          "GFLOP": round(v[3] / 1e9, 3)} for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])
                      if k not in host}
    out["host_wall"] = {"note": "calls and host wall-clock ms over the whole run: device allocations and frees "
                                "of the library, CSR rebuilds (allocation, row move and free)",
                        **{k: {"calls": prof[k][0], "ms": round(prof[k][1], 1)} for k in host if k in prof}}
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(w, ndirty, nmerged, args.steps)
        # vs_baseline stays null (no published number for this metric); the
        # GPU/CPU ratio is reported beside it, extrapolated like the CPU leg
        out["vs_cpu_baseline"] = {"ratio": round(value / out["cpu_baseline"]["value"], 2),
                                  "extrapolated": True}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
