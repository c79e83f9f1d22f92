"""Shared test helpers: small seeded graphs and the exact-PPR quantities used
as closed-form checks (mass conservation and the column-mass bound)."""
import numpy as np
import scipy.sparse as sp


def random_graph(n, m, seed, directed, no_self_loops=True):
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    if no_self_loops:
        keep = u != v
        u, v = u[keep], v[keep]
    return u, v


def transition(n, rowptr, col):
    deg = np.diff(rowptr).astype(np.float64)
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    w = 1.0 / deg[rows]
    P = sp.csr_matrix((w, (rows, col)), shape=(n, n))
    return P, deg


def reserve_mass(n, rowptr, col, alpha, iters=400):
    """m_v = sum_t pi_v(t): the part of v's PPR mass that ends in reserves
    under the push semantics (dangling vertices absorb without reserve):
    m = alpha * sum_k ((1-alpha) P)^k nd."""
    P, deg = transition(n, rowptr, col)
    nd = (deg > 0).astype(np.float64)
    m = np.zeros(n); x = alpha * nd
    for _ in range(iters):
        m += x
        x = (1 - alpha) * (P @ x)
        if np.abs(x).max() < 1e-17:
            break
    return m


def column_mass(n, rowptr, col, alpha, iters=400):
    """c_t = sum_v pi_v(t) (column sums of the exact PPR matrix):
    c = alpha * nd * sum_k ((1-alpha) P^T)^k 1."""
    P, deg = transition(n, rowptr, col)
    nd = (deg > 0).astype(np.float64)
    PT = P.T.tocsr()
    c = np.zeros(n); x = np.ones(n)
    for _ in range(iters):
        c += x
        x = (1 - alpha) * (PT @ x)
        if np.abs(x).max() < 1e-17:
            break
    return alpha * nd * c
