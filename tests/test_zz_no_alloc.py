"""The no-allocation frPCA entry point (tsvd_frpca_csr_noalloc) through the C
ABI: device-resident CSR in, device U, S, V out, every temporary in the
library's static per-GPU workspace.  Its results must equal the existing
entry point tsvd_frpca_csr (same seed, so the same sketch) and the oracle's
FP64 frPCA (frpca.c:239-326) with that sketch, on two calls in a row, the
second on a smaller matrix (the workspace keeps the first call's contents)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def ppr_like_block(m, n, density, seed):
    """Sparse non-negative block with the transform's value range
    (log10(1 + x / rmax) of reserves) and a decaying spectrum."""
    rng = np.random.Generator(np.random.PCG64(seed))
    A = sp.random(m, n, density=density, format="csr", random_state=rng,
                  data_rvs=lambda c: np.log10(1.0 + rng.lognormal(3.0, 2.0, c)))
    # a few dense-ish rows and columns, as hub vertices give
    A = A + sp.csr_matrix((np.abs(rng.standard_normal((m, 1))) @ np.abs(rng.standard_normal((1, n)))) *
                          (rng.random((m, n)) < density))
    return sp.csr_matrix(A)


def same_up_to_sign(X, Y, tol):
    for i in range(X.shape[1]):
        if min(np.abs(X[:, i] - Y[:, i]).max(), np.abs(X[:, i] + Y[:, i]).max()) > tol:
            return False
    return True


def test_noalloc_equals_the_allocating_entry_point_on_two_calls(capi):
    cases = [(2000, 12000, 0.01, 128, 6, 11),    # rank 128 as the configurations use
             (250, 700, 0.05, 16, 4, 12)]        # second call: smaller
    for m, n, dens, k, q, seed in cases:
        A = ppr_like_block(m, n, dens, seed)
        print(f"call m={m} n={n} nnz={A.nnz} k={k} q={q}", flush=True)
        U0, S0, V0 = capi.frpca_csr(A, k, q, seed)
        U1, S1, V1 = capi.frpca_csr_noalloc(A, k, q, seed)
        assert np.all(np.diff(S1) >= -1e-12), "S ascending like the reference"
        assert np.allclose(S1, S0, rtol=1e-9, atol=1e-12), np.abs(S1 - S0).max()
        # the leading factors agree to rounding; every factor obeys A^T U = V S
        top = np.flatnonzero(S0 >= 1e-2 * S0.max())
        gaps = np.abs(np.diff(S0)) / S0.max()
        sep = [i for i in top if (i == 0 or gaps[i - 1] > 1e-6) and (i == k - 1 or gaps[i] > 1e-6)]
        assert same_up_to_sign(U1[:, sep], U0[:, sep], 1e-6)
        assert same_up_to_sign(V1[:, sep], V0[:, sep], 1e-6)
        assert np.abs(A.T @ U1 - V1 * S1).max() <= 1e-9 * S1.max()
        assert np.allclose(U1.T @ U1, np.eye(k), atol=1e-9)
        # and the oracle with the same sketch
        _, So, _ = orc.frpca(A.toarray(), k, q, omega=orc.sketch(seed, n, k + 5))
        assert np.allclose(S1, So, rtol=1e-9, atol=1e-12)


def test_noalloc_rejects_bad_device_csr(capi):
    A = ppr_like_block(60, 90, 0.1, 3)
    bad = A.copy()
    bad.indices[3] = 90          # 1-based 91: outside [1, n]
    with pytest.raises(capi.TreeSVDError, match="column index"):
        capi.frpca_csr_noalloc(bad, 8, 6)
    # a good call afterwards still works
    U, S, V = capi.frpca_csr_noalloc(A, 8, 6, 5)
    _, So, _ = orc.frpca(A.toarray(), 8, 6, omega=orc.sketch(5, 90, 13))
    assert np.allclose(S, So, rtol=1e-9, atol=1e-12)
