"""The oracle -- TEST INFRASTRUCTURE ONLY.

A CPU restatement of the reference's update->embedding path, used by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  The
product library never imports it.

* Graph ingest, dynamic PPR update and FIFO forward push: plain C in
  oracle/treesvd_oracle.c (ctypes below), built by oracle/Makefile.
* Block map, elementwise transform, lazy dirty test, frPCA, tree merge and the
  root embedding: numpy FP64 below, each citing the reference lines.
* The n >= 1e7-node path (Pipeline(sparse_state=True)): the edge-order update
  and the positive-only storage in C, blocks from p + p^T with the
  30-interval row cut here.

Pinning of each function (the tests that check it) is listed in its docstring;
`exact_ppr` and the textbook numpy SVD are the closed forms used there.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double
        L.orc_graph_new.restype = vp; L.orc_graph_new.argtypes = [i32, i32]
        L.orc_graph_free.argtypes = [vp]
        L.orc_graph_append.restype = i32; L.orc_graph_append.argtypes = [vp, vp, vp, i64]
        L.orc_graph_degrees.argtypes = [vp, i32, vp, vp]
        L.orc_graph_csr.restype = i64; L.orc_graph_csr.argtypes = [vp, i32, vp, vp]
        L.orc_read_edge_file.restype = i64
        L.orc_read_edge_file.argtypes = [ctypes.c_char_p, i32, vp, vp, i64]
        L.orc_ppr_new.restype = vp
        L.orc_ppr_new.argtypes = [i32, i32, vp, i32, f64, f64, i32, i32]
        L.orc_ppr_free.argtypes = [vp]
        for f in ("orc_ppr_p", "orc_ppr_r"):
            getattr(L, f).restype = ctypes.POINTER(f64); getattr(L, f).argtypes = [vp, i32]
        L.orc_ppr_M.restype = ctypes.POINTER(f64); L.orc_ppr_M.argtypes = [vp]
        L.orc_ppr_init_static.argtypes = [vp, i32]
        L.orc_ppr_refresh_dynamic.argtypes = [vp, vp, i32, i32]
        L.orc_ppr_push.argtypes = [vp, vp, i32, i32]
        L.orc_ppr_counts.argtypes = [vp, vp, vp]
        L.orc_ppr_set_two_sided.argtypes = [vp, i32]
        L.orc_ppr_refresh_edges.argtypes = [vp, vp, i32, vp, vp, i64, i32]
        L.orc_ppr_drop_negative.argtypes = [vp, i32]
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# --------------------------------------------------------------------- graph
class Graph:
    """UGraph / Graph of Graph_dynamic.h:62-476, 495-885 (insert-only rows)."""

    def __init__(self, n: int, directed: bool):
        self.n, self.directed = n, directed
        self.h = lib().orc_graph_new(n, int(directed))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_graph_free(self.h)
            self.h = None

    def append(self, src, dst):
        s = np.ascontiguousarray(src, dtype=np.int32)
        d = np.ascontiguousarray(dst, dtype=np.int32)
        rc = lib().orc_graph_append(self.h, _ptr(s), _ptr(d), s.size)
        if rc != 0:
            raise ValueError("edge id out of range")

    def degrees(self, d=0):
        deg = np.zeros(self.n, np.int32); former = np.zeros(self.n, np.int32)
        lib().orc_graph_degrees(self.h, d, _ptr(deg), _ptr(former))
        return deg, former

    def csr(self, d=0):
        rp = np.zeros(self.n + 1, np.int64)
        m = lib().orc_graph_csr(self.h, d, _ptr(rp), None)
        col = np.zeros(max(m, 1), np.int32)
        lib().orc_graph_csr(self.h, d, _ptr(rp), _ptr(col))
        return rp, col[:m]


def read_edge_file(path: str, dup_last: bool = True):
    """The reference's snapshot-file loop (Graph_dynamic.h:127-130, 539-542)."""
    L = lib()
    cnt = L.orc_read_edge_file(path.encode(), int(dup_last), None, None, 0)
    if cnt < 0:
        raise IOError(f"cannot read {path} ({cnt})")
    s = np.zeros(max(cnt, 1), np.int32); d = np.zeros(max(cnt, 1), np.int32)
    L.orc_read_edge_file(path.encode(), int(dup_last), _ptr(s), _ptr(d), cnt)
    return s[:cnt], d[:cnt]


# ----------------------------------------------------------------------- PPR
class PPR:
    """Per-source reserves/residues and the FP64 matrix accumulator of the
    drivers (Tree_u_NC.cpp:459-489, d_row_tree_mkl::dense_mat_mapping)."""

    def __init__(self, n, sources, directed, alpha, rmax, fp32=True, quirk_int_ratio=False, two_sided=False):
        self.n, self.directed = n, directed
        self.sources = np.ascontiguousarray(sources, dtype=np.int32)
        self.nsrc = self.sources.size
        self.alpha, self.rmax = alpha, rmax
        self.h = lib().orc_ppr_new(n, self.nsrc, _ptr(self.sources), int(directed),
                                   alpha, rmax, int(fp32), int(quirk_int_ratio))
        if not self.h:
            raise MemoryError("oracle PPR state")
        lib().orc_ppr_set_two_sided(self.h, int(two_sided))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ppr_free(self.h)
            self.h = None

    def dirs(self):
        return (0, 1) if self.directed else (0,)

    def init_static(self):
        for d in self.dirs():
            lib().orc_ppr_init_static(self.h, d)

    def refresh_dynamic(self, g: Graph, nthreads=1):
        for d in self.dirs():
            lib().orc_ppr_refresh_dynamic(self.h, g.h, d, nthreads)

    def push(self, g: Graph, nthreads=1):
        for d in self.dirs():
            lib().orc_ppr_push(self.h, g.h, d, nthreads)

    def refresh_edges(self, g: Graph, src, dst, nthreads=1):
        """The n >= 1e7-node path's insertion update, edge by edge in batch
        order (ppr:6504-6533, 6675-6704; Q8 included)."""
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        for d in self.dirs():
            lib().orc_ppr_refresh_edges(self.h, g.h, d, _ptr(src), _ptr(dst), src.size, nthreads)

    def drop_negative(self):
        """Keep only positive entries, as the pair lists do (ppr:6598-6626)."""
        for d in self.dirs():
            lib().orc_ppr_drop_negative(self.h, d)

    def counts(self):
        """(pushes, relaxations) since the last call."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        lib().orc_ppr_counts(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def p(self, d=0):
        return np.ctypeslib.as_array(lib().orc_ppr_p(self.h, d), shape=(self.nsrc, self.n))

    def r(self, d=0):
        return np.ctypeslib.as_array(lib().orc_ppr_r(self.h, d), shape=(self.nsrc, self.n))

    def M(self):
        return np.ctypeslib.as_array(lib().orc_ppr_M(self.h), shape=(self.nsrc, self.n))


def exact_ppr(n, rowptr, col, alpha, sources, directed):
    """Exact FP64 PPR rows under the reference's push semantics (closed form
    for tiny graphs): pi_s = alpha * h_s (restricted to non-dangling nodes),
    h_s = e_s + (1-alpha) h_s D^+ A, where a dangling (degree-0) node absorbs
    its mass and gets no reserve (directed, ppr:2971-2976).  Undirected
    degree-0 sources keep pi = alpha (ppr:1431-1441)."""
    deg = np.diff(rowptr).astype(np.float64)
    P = np.zeros((n, n))
    for v in range(n):
        for j in range(rowptr[v], rowptr[v + 1]):
            P[v, col[j]] += 1.0 / deg[v]
    I = np.eye(n)
    H = np.linalg.solve((I - (1 - alpha) * P).T, I[:, sources]).T   # rows h_s
    nd = (deg > 0).astype(np.float64)
    pi = alpha * H * nd
    if not directed:
        for k, s in enumerate(sources):
            if deg[s] == 0:
                pi[k, s] = alpha
    return pi, P


# ---------------------------------------------------------- blocks and tree
def block_map(n: int, nparts: int):
    """Column blocks (Tree_u_NC.cpp:396-432): common = n // nparts, block id
    min(v // common, nparts-1), the last block takes the remainder.
    Returns (starts, widths)."""
    common = n // nparts
    starts = np.arange(nparts, dtype=np.int64) * common
    widths = np.full(nparts, common, dtype=np.int64)
    widths[-1] = n - (nparts - 1) * common
    return starts, widths


def transform(X, rmax, mode="NC"):
    """Log_sparse_matrix_entries (ppr:2527-2561, NC: log10(x/rmax) if x>rmax
    else 0) and Log_sparse_matrix_entries_LP (ppr:2598-2638, log10(1+x/rmax)
    on every stored (non-zero) entry)."""
    X = np.asarray(X, dtype=np.float64)
    if mode == "NC":
        out = np.zeros_like(X)
        m = X > rmax
        out[m] = np.log10(X[m] / rmax)
        return out
    out = np.zeros_like(X)
    m = X != 0
    out[m] = np.log10(1.0 + X[m] / rmax)
    return out


def dirty_test(A_new, A_old, resid, delta):
    """Lazy-update test of Log_sparse_matrix_entries_with_norm_computation
    (ppr:2699-2726, paper Eqn 2): clean iff resid + ||A_new - A_old||_F <
    delta*sqrt(2)*||A_new||_F.  Returns (dirty, ||A_new||, ||A_new-A_old||)."""
    an = float(np.linalg.norm(A_new))
    en = float(np.linalg.norm(A_new - A_old))
    return not (resid + en < delta * np.sqrt(2.0) * an), an, en


def tree_levels(nparts: int, fanout: int):
    """Level sizes from the leaves up.  For nparts == fanout**(L-1) this is the
    reference's heap tree (ppr:225-245; children p*f+1..p*f+f); otherwise
    parent i of a level holds children [i*f, min((i+1)*f, size)) of the level
    below (ragged tree)."""
    sizes = [nparts]
    while sizes[-1] > 1:
        sizes.append((sizes[-1] + fanout - 1) // fanout)
    return sizes


def lu_fraction(A):
    """LUfraction (frpca.c:4-41): [L, ~] = lu(A) with the row swaps undone,
    i.e. L = P @ L_unit where A = P L U."""
    import scipy.linalg
    P, L, _ = scipy.linalg.lu(A)
    return P @ L


def eig_svd(A, eps=1e-4):
    """eigSVD (frpca.c:99-136): V S^2 V^T = A^T A (dsyevd, ascending);
    S = sqrt(max(lambda, 0)); columns with S <= eps are zeroed; U = A V / S."""
    G = A.T @ A
    lam, V = np.linalg.eigh(G)
    S = np.sqrt(np.maximum(lam, 0.0))
    V1 = np.where(S > eps, V / np.where(S > eps, S, 1.0), 0.0)
    S = np.where(S > eps, S, 0.0)
    U = A @ V1
    return U, S, V


M64 = (1 << 64) - 1


def _mix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def sketch(seed, rows, cols, fp32=False):
    """The counter-based Gaussian sketch used on the device (the reference
    draws it from MKL VSL, mkl.c:479-512): entry (i, c) is Box-Muller of a
    splitmix64 hash of seed and the column-major index c*rows + i.  Used so
    that device and oracle factor with the same sketch."""
    with np.errstate(over="ignore"):
        t = np.arange(rows * cols, dtype=np.uint64)
        h = _mix64(np.uint64(seed & M64) ^ _mix64(t))
    u1 = ((h >> np.uint64(40)).astype(np.float64) + 1.0) / 16777217.0
    u2 = (h & np.uint64(0xFFFFFF)).astype(np.float64) / 16777216.0
    x = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)
    x = x.reshape(cols, rows).T
    return x.astype(np.float32).astype(np.float64) if fp32 else x


def leaf_seed(seed, j):
    """Sketch seed of column block j (matches the device).  It does not depend
    on the snapshot, so a block's factor does not depend on how insertions
    were grouped into snapshots."""
    seed, j = int(seed), int(j)
    return (seed * 0x9E3779B97F4A7C15 + j * 0xD1B54A32D192ED03 + 1) & M64


def frpca(A, k, q, rng=None, s=5, omega=None):
    """frPCA (frpca.c:239-326) in FP64.  The sketch is `omega` when given
    (n x l), else drawn from `rng` (the reference seeds VSL with time(NULL),
    mkl.c:495).  Returns U (m x k), S (k, ascending as the reference),
    V (n x k)."""
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    l = k + s
    if q % 2 == 0:
        Qt = omega if omega is not None else rng.standard_normal((n, l)).astype(np.float32).astype(np.float64)
        Q = A @ Qt
        if q == 2:
            Q, SS, VV = eig_svd(Q)
        else:
            Q = lu_fraction(Q)
    else:
        Q = rng.standard_normal((m, l)).astype(np.float32).astype(np.float64)
    niter = (q - 1) // 2
    for i in range(1, niter + 1):
        Qt = A.T @ Q
        Q = A @ Qt
        if i == niter:
            Q, SS, VV = eig_svd(Q)
        else:
            Q = lu_fraction(Q)
    Qt = A.T @ Q
    UU, SS, VV = eig_svd(Qt)
    inds = np.arange(s, k + s)
    V = UU[:, inds]
    S = SS[inds]
    U = Q @ VV[:, inds]
    return U, S, V


def truncated_svd(C, k):
    """truncated_singular_value_decomposition (mkl.c:1259-1322): dgesvd('S',
    'S'), top-k (descending)."""
    U, S, Vt = np.linalg.svd(C, full_matrices=False)
    return U[:, :k], S[:k], Vt[:k]


def merge(children, k):
    """dense_sub_svd_function (Tree_u_NC.cpp:238-262) over C = [child_1 | ...]
    (append_matrices_horizontally, mkl.c:1024-1054): parent = U_k diag(S_k)."""
    C = np.concatenate(children, axis=1)
    U, S, _ = truncated_svd(C, k)
    return U * S, U, S


def lp_precision(scores, positive, m):
    """Link-prediction precision of Tree_u_LP.cpp:1092-1110: the m pairs with
    the highest scores (nth_element + sort), the share of them that are
    positive pairs.  Ties are broken by input order."""
    order = np.argsort(-np.asarray(scores, dtype=np.float64), kind="stable")[:m]
    return float(np.count_nonzero(np.asarray(positive)[order])) / m


def root_embedding(children, k, mode="NC"):
    """Root of the merge loop: NC E = U_k diag(sqrt|S_k|) (Tree_u_NC.cpp:
    752-786); LP E = U_k (Tree_u_LP.cpp:907-912)."""
    _, U, S = merge(children, k)
    return U * np.sqrt(np.abs(S)) if mode == "NC" else U


class Pipeline:
    """The per-snapshot driver of Tree_u_NC.cpp:507-866 / Tree_d_NC.cpp:
    611-1132 (and the LP root, Tree_u_LP.cpp:664-925) in FP64 numpy + the C
    push.  dyn_start: first snapshot with dynamic updates; lazy_start: first
    snapshot with the dirty test (residual norms from lazy_start-1).

    Each dirty block is factored over its active columns (vertices with an
    edge, or sources -- every other column is zero) with the device's sketch
    for (seed, snapshot, block), so that device and oracle differ only by
    their push order and arithmetic."""

    def __init__(self, n, directed, sources, alpha, rmax, nparts, fanout, k, q,
                 mode="NC", delta=0.65, dyn_start=2, lazy_start=2, fp32=True,
                 seed=0, nthreads=os.cpu_count() or 1, two_sided=False, sparse_state=False):
        self.g = Graph(n, directed)
        self.ppr = PPR(n, sources, directed, alpha, rmax, fp32=fp32, two_sided=two_sided)
        self.sources = np.asarray(sources)
        self.n, self.directed, self.rmax = n, directed, rmax
        self.nparts, self.fanout, self.k, self.q = nparts, fanout, k, q
        self.mode, self.delta = mode, delta
        self.dyn_start, self.lazy_start = dyn_start, lazy_start
        self.starts, self.widths = block_map(n, nparts)
        self.levels = tree_levels(nparts, fanout)
        nsrc = self.sources.size
        self.A_old = [np.zeros((nsrc, w)) for w in self.widths]
        self.resid = np.zeros(nparts)
        # U*S per tree node, zero until first factored (matrix_new is calloc)
        self.nodes = [[np.zeros((nsrc, k)) for _ in range(sz)] for sz in self.levels]
        self.seed = seed
        self.nthreads = nthreads
        self.E = None
        self.root_sigma = None
        self.V = None                             # LP right embedding (n x k)
        self.U_of = None
        self.it = 0
        self.last_dirty = []
        # the n >= 1e7-node path (Tree_d_LP.cpp:1377-2024): directed LP only
        self.sparse_state = sparse_state
        if sparse_state and not (directed and mode == "LP" and not two_sided):
            raise ValueError("sparse_state is the directed LP path with the positive-side threshold")
        self.keep = 30 * (nsrc // 30)

    def matrix(self):
        if self.sparse_state:
            # ppr:6898-6960: forward and transpose reserve lists summed, rows
            # inserted in 30 intervals of floor(|S|/30) (the rest stay empty, Q11)
            X = self.ppr.p(0) + self.ppr.p(1)
            X[self.keep:] = 0.0
            return X
        return self.ppr.M()

    def active(self):
        deg = self.g.degrees(0)[0].astype(np.int64)
        if self.directed:
            deg = deg + self.g.degrees(1)[0]
        act = deg > 0
        act[self.sources] = True
        return act

    def snapshot(self, src, dst):
        self.it += 1
        it = self.it
        self.g.append(src, dst)
        # Tree_u_NC.cpp:516-518 (sentinel 200 in the n >= 1e7-node path, Tree_d_LP.cpp:1574)
        if self.lazy_start != (200 if self.sparse_state else 100) and it < self.lazy_start - 1:
            return self.E
        if it < self.dyn_start:
            self.ppr.init_static()
        elif self.sparse_state:
            self.ppr.refresh_edges(self.g, src, dst, self.nthreads)
        else:
            self.ppr.refresh_dynamic(self.g, self.nthreads)
        self.ppr.push(self.g, self.nthreads)
        if self.sparse_state:
            self.ppr.drop_negative()
        M = self.matrix()
        act = self.active()
        dirty = []
        for j in range(self.nparts):
            c0, w = self.starts[j], self.widths[j]
            A = transform(M[:, c0:c0 + w], self.rmax, self.mode)
            d = True
            # the n >= 1e7-node path tests from dyn_start on (Tree_d_LP.cpp:1658)
            if it >= (self.dyn_start if self.sparse_state else self.lazy_start):
                d, _, _ = dirty_test(A, self.A_old[j], self.resid[j], self.delta)
            if d:
                self.A_old[j] = A
                if np.count_nonzero(A):
                    dirty.append(j)
        self.last_dirty = dirty
        if not dirty:
            return self.E                       # Tree_u_NC.cpp:713-721
        for j in dirty:
            c0, w = self.starts[j], self.widths[j]
            cols = np.flatnonzero(act[c0:c0 + w])
            Ac = self.A_old[j][:, cols]
            om = sketch(leaf_seed(self.seed, j), cols.size, self.k + 5)
            U, S, V = frpca(Ac, self.k, self.q, omega=om)
            self.nodes[0][j] = U * S
            if it >= self.lazy_start - 1:        # Tree_u_NC.cpp:158-204
                self.resid[j] = np.linalg.norm((U * S) @ V.T - Ac)
        touched = sorted({j // self.fanout for j in dirty})
        top = len(self.levels) - 1
        for lvl in range(1, top + 1):
            below = self.nodes[lvl - 1]
            for p in touched:
                ch = below[p * self.fanout:min((p + 1) * self.fanout, len(below))]
                if lvl == top:
                    self.E = root_embedding(ch, self.k, self.mode)
                    self.root_sigma = merge(ch, self.k)[2]
                else:
                    self.nodes[lvl][p] = merge(ch, self.k)[0]
            touched = sorted({p // self.fanout for p in touched})
        if top == 0:                              # a single block: its own root
            _, U, S = merge([self.nodes[0][0]], self.k)
            self.E = U * np.sqrt(np.abs(S)) if self.mode == "NC" else U
            self.root_sigma = S
        if self.mode == "LP":                     # Tree_u_LP.cpp:1040-1045, 291-404
            if self.V is None:
                self.V = np.zeros((self.n, self.k))
                self.U_of = [None] * self.nparts
            for j in dirty:
                c0, w = self.starts[j], self.widths[j]
                self.V[c0:c0 + w] = self.A_old[j].T @ self.E
                self.U_of[j] = self.E
        return self.E

    def lp_scores(self, src, dst):
        """Tree_u_LP.cpp:1047-1090: score(i, v) = U_b[row(i)] . V[v] with U_b
        the root U that block(v)'s V was computed with."""
        row = {int(v): i for i, v in enumerate(self.sources)}
        common = self.n // self.nparts
        out = np.zeros(len(src))
        for e, (i, v) in enumerate(zip(src, dst)):
            b = min(int(v) // common, self.nparts - 1)
            U = self.U_of[b]
            out[e] = 0.0 if U is None else U[row[int(i)]] @ self.V[int(v)]
        return out
