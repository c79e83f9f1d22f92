"""frPCA through the drop-in C ABI (tsvd_frpca_csr / frPCA, the signature of
frpca.c:239): against the oracle's FP64 restatement run with the same
Gaussian sketch (span-equivalent, so singular values agree to rounding), and
against the textbook FP64 SVD (Eckart-Young optimality)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def make_matrix(m, n, seed, density=0.3, decay=0.85):
    rng = np.random.Generator(np.random.PCG64(seed))
    U, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
    V, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
    s = 10.0 * decay ** np.arange(min(m, n))
    A = (U * s) @ V.T
    A[rng.random(A.shape) > density] = 0.0
    return A


@pytest.mark.parametrize("m,n,k,q", [(120, 300, 16, 6), (300, 120, 20, 4), (64, 700, 8, 2), (200, 400, 30, 6)])
def test_dropin_matches_oracle_same_sketch(capi, m, n, k, q):
    A = make_matrix(m, n, m + n)
    seed = 1234 + m
    U, S, V = capi.frpca_csr(A, k, q, seed)
    Uo, So, Vo = orc.frpca(A, k, q, omega=orc.sketch(seed, n, k + 5))
    assert np.all(np.diff(S) >= -1e-12), "S must be ascending like the reference"
    assert np.allclose(S, So, rtol=1e-9, atol=1e-12)
    # same factors up to column signs
    for i in range(k):
        assert min(np.abs(U[:, i] - Uo[:, i]).max(), np.abs(U[:, i] + Uo[:, i]).max()) < 1e-7
        assert min(np.abs(V[:, i] - Vo[:, i]).max(), np.abs(V[:, i] + Vo[:, i]).max()) < 1e-7
    assert np.allclose(U.T @ U, np.eye(k), atol=1e-9)


def test_dropin_eckart_young(capi):
    A = make_matrix(250, 600, 5, density=0.5, decay=0.9)
    k, q = 20, 6
    U, S, V = capi.frpca_csr(A, k, q, 77)
    s_exact = np.linalg.svd(A, compute_uv=False)
    best = np.sqrt((s_exact[k:] ** 2).sum())
    err = np.linalg.norm(A - (U * S) @ V.T)
    assert err <= 1.02 * best
    # randomized SVD: the leading half of the spectrum is resolved to 1%
    rel = np.abs(np.sort(S)[::-1] - s_exact[:k]) / s_exact[:k]
    assert rel[:k // 2].max() < 1e-2, f"sigma errors {rel}"
    # residual identity used by the pipeline: ||A - U S V^T||^2 = ||A||^2 - sum S^2
    assert abs(err ** 2 - (np.linalg.norm(A) ** 2 - (S ** 2).sum())) < 1e-8 * np.linalg.norm(A) ** 2


def test_dropin_frPCA_symbol_and_zero_rows(capi):
    """The void frPCA(mat_csr*, mat**, mat**, mat**, int, int) entry point, on a
    matrix with empty rows (pointerB == pointerE)."""
    import ctypes
    A = make_matrix(80, 200, 9)
    A[::7] = 0.0
    Asp = sp.csr_matrix(A)
    m, n, k = 80, 200, 10
    vals = np.ascontiguousarray(Asp.data, np.float64)
    cols = np.ascontiguousarray(Asp.indices + 1, np.int32)
    pb = np.ascontiguousarray(Asp.indptr[:-1] + 1, np.int32)
    pe = np.ascontiguousarray(Asp.indptr[1:] + 1, np.int32)
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    csr = capi.MatCSR(Asp.nnz, m, n, P(vals, ctypes.c_double), P(cols, ctypes.c_int), P(pb, ctypes.c_int),
                      P(pe, ctypes.c_int))
    Ub, Sb, Vb = np.zeros(m * k), np.zeros(k), np.zeros(n * k)
    mu = capi.Mat(m, k, P(Ub, ctypes.c_double)); ms = capi.Mat(k, 1, P(Sb, ctypes.c_double))
    mv = capi.Mat(n, k, P(Vb, ctypes.c_double))
    pu, ps, pv = ctypes.pointer(mu), ctypes.pointer(ms), ctypes.pointer(mv)
    capi.lib().frPCA(ctypes.byref(csr), ctypes.byref(pu), ctypes.byref(ps), ctypes.byref(pv), k, 6)
    s_exact = np.linalg.svd(A, compute_uv=False)[:k]
    rel = np.abs(np.sort(Sb)[::-1] - s_exact) / s_exact
    assert rel[:k // 2].max() < 1e-2, f"sigma errors {rel}"


def test_dropin_rejects_bad_csr(capi):
    A = sp.csr_matrix(make_matrix(30, 40, 3))
    bad = A.copy()
    bad.indices[0] = 40   # out of range (becomes 41, 1-based)
    with pytest.raises(capi.TreeSVDError):
        capi.frpca_csr(bad, 5, 6)


def test_batched_dropin_equals_single_calls(capi):
    """tsvd_frpca_csr_batch over matrices of different shapes (the drivers'
    dirty blocks): each item equals tsvd_frpca_csr with seed + i, and the
    oracle with that item's sketch."""
    shapes = [(96, 300), (96, 157), (96, 41), (96, 520)]
    mats = [make_matrix(m, n, 40 + i) for i, (m, n) in enumerate(shapes)]
    k, q, seed = 12, 6, 900
    out = capi.frpca_csr_batch(mats, k, q, seed)
    for i, (A, (U, S, V)) in enumerate(zip(mats, out)):
        U1, S1, V1 = capi.frpca_csr(A, k, q, seed + i)
        assert np.allclose(S, S1, rtol=1e-12, atol=1e-14), f"item {i}"
        _, So, _ = orc.frpca(A, k, q, omega=orc.sketch(seed + i, A.shape[1], k + 5))
        assert np.allclose(S, So, rtol=1e-9, atol=1e-12), f"item {i} vs oracle"
        assert np.allclose((U * S) @ V.T, (U1 * S1) @ V1.T, atol=1e-9 * np.abs(A).max())


def test_batched_dropin_rejects_bad_arguments(capi):
    mats = [make_matrix(50, 80, 1), make_matrix(50, 80, 2)]
    with pytest.raises(capi.TreeSVDError):
        capi.frpca_csr_batch(mats, 8, 0, 1)       # q must be positive
    with pytest.raises(capi.TreeSVDError):
        capi.frpca_csr_batch([], 8, 6, 1)         # empty batch
