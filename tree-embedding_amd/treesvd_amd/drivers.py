"""The reference drivers' per-snapshot loop over their own files, on the
plain GPU method.

Tree_u_NC.cpp / Tree_d_NC.cpp / Tree_u_LP.cpp / Tree_d_LP.cpp read
DY_Dataset/<name>/config.txt (line 0: the subset file, then one edge file per
snapshot), run one update per snapshot file and write U_all_shots -- the
stacked per-snapshot subset embeddings, (snapshots-1)*|S| x d -- as a CSV
(Tree_u_NC.cpp:338-379, 507-866, 898-909).  run_config does the same through
the C ABI; argument names follow the drivers' argv (Tree_u_NC.cpp:301-322).
"""
from __future__ import annotations

import numpy as np

from . import capi


def run_config(config_path: str, out_csv: str | None, vertex_number: int, alpha: float = 0.5, pass_: int = 12,
               residuemax: float = 1e-6, nParts: int = 64, hierarchy_n: int = 8, dynamic_ppr_start_iter: int = 2,
               lazy_update_start_iter: int = 2, delta: float = 0.65, d: int = 128, directed: bool = False,
               mode: str = "NC", precision: int = capi.TSVD_FP32, dup_last: bool = True, seed: int = 0):
    """Runs every snapshot file named in config.txt and returns U_all_shots
    ((snapshots-1) * |S| x d, row-major); writes it to out_csv when given.
    A snapshot that the drivers skip (lazy_update_start_iter != 100 and
    iter < lazy_update_start_iter - 1, Tree_u_NC.cpp:516-518) leaves its rows
    zero here; the drivers leave them as MatrixXd allocated them
    (uninitialised)."""
    files = capi.read_lines(config_path)
    if len(files) < 2:
        raise ValueError(f"{config_path}: needs a subset file and at least one snapshot file")
    subset = capi.read_subset_file(files[0])
    pipe = capi.Pipeline(vertex_number, directed, subset, alpha, residuemax, nParts, hierarchy_n, d, pass_,
                         precision=precision,
                         transform=capi.TSVD_TRANSFORM_NC if mode == "NC" else capi.TSVD_TRANSFORM_LP,
                         delta=delta, dyn_start=dynamic_ppr_start_iter, lazy_start=lazy_update_start_iter,
                         seed=seed)
    U_all = np.zeros(((len(files) - 1) * subset.size, d))
    try:
        for it in range(1, len(files)):
            s, t = capi.read_edge_file(files[it], dup_last)
            pipe.snapshot(s, t)
            if lazy_update_start_iter != 100 and it < lazy_update_start_iter - 1:
                continue                         # ingested only
            U_all[(it - 1) * subset.size:it * subset.size] = pipe.embedding()
    finally:
        pipe.close()
    if out_csv:
        capi.write_embedding_csv(out_csv, U_all)
    return U_all
