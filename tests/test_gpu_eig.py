"""The batched FP64 symmetric eigensolver (tsvd_eig_batched, the device's
replacement for eigSVD's LAPACKE_dsyevd, frpca.c:99-136, and for the merge
SVDs, matrix_vector_functions_intel_mkl.c:1259-1322) against LAPACK's dsyevd
through numpy.linalg.eigh.

Bars, for every matrix of every batch: eigenvalues within 1e-13 ||A||_2 of
LAPACK's, residual ||A V - V diag(lam)||_F <= 1e-12 ||A||_2, and
||V^T V - I||_max <= 1e-12.  The shapes are those the pipeline solves (Grams of
l = k + 5 <= 136 columns in LDS, merges of f * k = 512 in global memory) and
the edges around them: n = 1 and 2, the LDS limit 136/137, sizes whose halves
split unevenly (97, 129, 385), 1024.  The spectra are those of the pipeline's Grams (graded, rank-
deficient, zero-padded blocks as ragged trees give) and the divide and
conquer's deflation cases (exact clusters, identity, zero, diagonal) and an
indefinite matrix, all in one batch per size, so that batch slots are not
confused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spectra(n, rng):
    mats = {}
    X = rng.standard_normal((max(3 * n, 8), n)) * 0.93 ** np.arange(n)
    mats["graded gram"] = X.T @ X
    r = max(n // 3, 1)
    Y = rng.standard_normal((r, n))
    mats["rank n/3 gram"] = Y.T @ Y
    G = np.zeros((n, n))
    h = max(n // 2, 1)
    Z = rng.standard_normal((2 * h, h))
    G[:h, :h] = Z.T @ Z
    mats["zero-padded gram"] = G
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.resize([5.0, 2.0, 1.0, 1e-9], n)
    mats["exact clusters"] = (Q * lam) @ Q.T
    mats["identity"] = 3.0 * np.eye(n)
    mats["zero"] = np.zeros((n, n))
    mats["diagonal"] = np.diag(rng.standard_normal(n))
    M = rng.standard_normal((n, n))
    mats["indefinite"] = M + M.T
    return mats


def check_batch(capi, mats, top=0):
    A = np.stack(list(mats.values()))
    lam, V, info = capi.eig_batched(A, top)
    assert info == 0
    n = A.shape[1]
    worst = {}
    for b, name in enumerate(mats):
        M = A[b]
        ref = np.linalg.eigvalsh(M)
        scale = max(np.linalg.norm(M, 2), 1e-300)
        cols = slice(0, n) if top <= 0 else slice(n - top, n)
        Vb = V[b][:, cols]
        lb = lam[b]
        e_lam = np.max(np.abs(lb - ref)) / scale
        e_res = np.linalg.norm(M @ Vb - Vb * lb[cols]) / scale if scale > 1e-300 else np.linalg.norm(Vb.T @ Vb - np.eye(Vb.shape[1]))
        e_orth = np.max(np.abs(Vb.T @ Vb - np.eye(Vb.shape[1])))
        worst[name] = (e_lam, e_res, e_orth)
        assert np.all(np.diff(lb) >= 0), name
        assert e_lam <= 1e-13, (name, n, e_lam)
        assert e_res <= 1e-12, (name, n, e_res)
        assert e_orth <= 1e-12, (name, n, e_orth)
    return lam, V, worst


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 96, 97, 129, 133, 136, 137, 200, 385, 512, 1024])
def test_eig_against_lapack(capi, n):
    rng = np.random.default_rng(n)
    _, _, worst = check_batch(capi, spectra(n, rng))
    print(f"n={n}: " + ", ".join(f"{k} {v[0]:.1e}/{v[1]:.1e}/{v[2]:.1e}" for k, v in worst.items()))


@pytest.mark.parametrize("n, top", [(133, 128), (512, 128), (96, 20), (300, 1)])
def test_top_eigenvectors_match_the_full_solve(capi, n, top):
    """top < n: the last merge forms only the top eigenvectors; they and all
    eigenvalues equal the full solve's (vectors up to sign) to rounding: the
    cross-wave sums of the tridiagonalization's products are added in
    arrival order, so two runs agree to a few ulps, not bit for bit."""
    rng = np.random.default_rng(100 + n)
    mats = spectra(n, rng)
    keep = {k: mats[k] for k in ("graded gram", "rank n/3 gram", "indefinite")}
    lam_t, V_t, _ = check_batch(capi, keep, top)
    lam_f, V_f, _ = check_batch(capi, keep, 0)
    for b, M in enumerate(keep.values()):
        assert np.max(np.abs(lam_t[b] - lam_f[b])) <= 1e-13 * max(np.linalg.norm(M, 2), 1e-300)
    # vectors are unique (up to sign) only for separated eigenvalues: the rank
    # n/3 Gram's zero eigenvalue is a cluster of about 2n/3
    for b, M in enumerate(keep.values()):
        scale = np.linalg.norm(M, 2)
        lam = lam_f[b]
        gap = np.minimum(np.diff(lam, prepend=-np.inf), np.diff(lam, append=np.inf))
        sep = np.where(gap[n - top:] > 1e-6 * scale)[0]
        a, f = V_t[b][:, n - top:][:, sep], V_f[b][:, n - top:][:, sep]
        sgn = np.sign(np.sum(a * f, axis=0))
        assert np.max(np.abs(a - f * sgn), initial=0.0) < 1e-8


def test_batch_of_pipeline_grams(capi):
    """A batch as the frPCA pass solves it: 143 Grams of 133 columns."""
    rng = np.random.default_rng(7)
    A = []
    for i in range(143):
        X = rng.standard_normal((400, 133)) * (0.9 + 0.09 * rng.random()) ** np.arange(133)
        A.append(X.T @ X)
    A = np.stack(A)
    lam, V, info = capi.eig_batched(A)
    assert info == 0
    for b in range(0, 143, 13):
        ref = np.linalg.eigvalsh(A[b])
        scale = np.linalg.norm(A[b], 2)
        assert np.max(np.abs(lam[b] - ref)) <= 1e-13 * scale
        assert np.linalg.norm(A[b] @ V[b] - V[b] * lam[b]) <= 1e-12 * scale


def test_eig_rejects_bad_sizes(capi):
    with pytest.raises(capi.TreeSVDError):
        capi.eig_batched(np.zeros((1, 1025, 1025)))
