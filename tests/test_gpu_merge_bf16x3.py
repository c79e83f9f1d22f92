"""bf16x3 merge products (tsvd_config.merge_bf16x3: the merge GEMMs C W_k of
Tree_u_NC.cpp:238-262 on bf16 MFMA with FP32 accumulation, each FP32
operand split into bf16 hi + lo) against the FP64 spectrum of each merge's
own children on the device.

For every internal node P = U_k Sigma_k of C = [children] and for the root:
the top-16 column norms of P are the top-16 square roots of eigvalsh(C^T C)
(FP64, LAPACK) within 1e-3 relative, and the merge residual (column norms,
orthogonality, C C^T P = P Sigma^2, relative to sigma_max) is below 1e-3.
The products keep about 16 significant bits; the Grams and eigensolves stay
FP64.  Trees: ragged fan-out 4 at rank 128 (the 144-column MFMA tile) and at
rank 16 (the 80-column tile), and the same pipeline with FP64 accumulation
as the control."""
import numpy as np
import pytest

from oracle import oracle as orc
from test_gpu_configs import merge_error, shape_workload, tree_nodes

pytestmark = pytest.mark.gpu


def spectrum_error(P, children, top=16):
    C = np.concatenate(children, axis=1)
    lam = np.linalg.eigvalsh(C.T @ C)[::-1][:top]
    sig = np.sqrt(np.clip(lam, 0.0, None))
    got = np.linalg.norm(P, axis=0)[:top]
    keep = sig > 1e-12 * max(sig[0], 1e-300)
    return float(np.max(np.abs(got[keep] - sig[keep]) / sig[keep], initial=0.0))


def check(pipe, nparts, fanout, k):
    levels, nodes = tree_nodes(pipe, nparts, fanout)
    worst_s = worst_m = 0.0
    for lvl in range(1, len(levels) - 1):
        for p_ in range(levels[lvl]):
            ch = nodes[lvl - 1][p_ * fanout:min((p_ + 1) * fanout, levels[lvl - 1])]
            worst_s = max(worst_s, spectrum_error(nodes[lvl][p_], ch))
            worst_m = max(worst_m, merge_error(nodes[lvl][p_], ch))
    top = len(levels) - 1
    sig = pipe.root_sigma()
    P = pipe.embedding() * np.sqrt(np.abs(sig))[None, :]
    worst_s = max(worst_s, spectrum_error(P, nodes[top - 1]))
    worst_m = max(worst_m, merge_error(P, nodes[top - 1]))
    return worst_s, worst_m


@pytest.mark.parametrize("n,m,nparts,k", [(1 << 15, 400_000, 40, 128), (1 << 15, 300_000, 70, 16)],
                         ids=["40-leaves-rank128", "70-leaves-rank16"])
def test_bf16x3_merges_against_the_fp64_spectrum(capi, n, m, nparts, k):
    n, snaps, sub = shape_workload(False, n, m, 256, 8)
    res = {}
    for bf3 in (True, False):
        pipe = capi.Pipeline(n, False, sub, 0.2, 1e-6, nparts, 4, k, 6, precision=capi.TSVD_FP32,
                             transform=capi.TSVD_TRANSFORM_NC, delta=0.3, seed=5, merge_bf16x3=bf3)
        try:
            for s, d in snaps:
                pipe.snapshot(s, d)
                res[bf3] = check(pipe, nparts, 4, k)
                print(f"merge_bf16x3={bf3}: top-16 sigma error {res[bf3][0]:.2e}, merge residual {res[bf3][1]:.2e}",
                      flush=True)
                assert res[bf3][0] <= 1e-3 and res[bf3][1] <= 1e-3
        finally:
            pipe.close()
    # FP64 accumulation is the tighter control; bf16x3 stays within its bar above
    assert res[False][1] <= 2e-4
