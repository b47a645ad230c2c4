"""Seeded synthetic matrices with exactly known null spaces.

Each generator returns a :class:`CsrMatrix` (canonical CSR: int32 row
pointers and column indices, columns strictly ascending within a row, FP64
values) and, where it is cheap, the exact kernel.

Random streams: ``numpy.random.Generator(PCG64(SeedSequence(seed,
spawn_key=(stream,))))`` with one documented ``stream`` id per purpose, so a
matrix depends only on (generator, sizes, seed).

Workloads (sizes are the defaults of the callers, not fixed here):
  * shifted difference (A x)_i = x_{i+N} - c x_i  (examples/quickstart.cpp:25-39)
  * oriented incidence of disjoint random 4-regular graphs, each the union of
    two random Hamiltonian cycles
  * weighted incidence of disjoint 4-neighbour grids, weights 10**(-3U)
  * A = [R | R G] with R diagonally dominant sparse, kernel span [-G; I]
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class CsrMatrix:
    m: int
    n: int
    rowptr: np.ndarray  # int32 (m+1)
    col: np.ndarray     # int32 (nnz)
    val: np.ndarray     # float64 (nnz)

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def fro(self):
        return float(np.sqrt(np.sum(self.val * self.val)))

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.m, self.n))


def _rng(seed, stream):
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence(seed, spawn_key=(stream,))))


def _from_two_per_row(m, n, ca, va, cb, vb):
    """Rows with exactly two entries (ca, va), (cb, vb), ca != cb."""
    lo = np.minimum(ca, cb)
    hi = np.maximum(ca, cb)
    vlo = np.where(ca < cb, va, vb)
    vhi = np.where(ca < cb, vb, va)
    col = np.empty(2 * m, np.int32)
    val = np.empty(2 * m, np.float64)
    col[0::2] = lo
    col[1::2] = hi
    val[0::2] = vlo
    val[1::2] = vhi
    rowptr = np.arange(0, 2 * m + 1, 2, dtype=np.int64).astype(np.int32)
    return CsrMatrix(m, n, rowptr, col, val)


def shifted_difference(n, N, c):
    """(n-N) x n, row i: -c at column i, +1 at column i+N."""
    m = n - N
    i = np.arange(m, dtype=np.int64)
    return _from_two_per_row(m, n, i, np.full(m, -c), i + N, np.ones(m))


def shifted_difference_kernel(n, N, c):
    """Orthonormal basis (n x N) of ker: (v_j)_{j+kN} = c^k (README.md:24-28)."""
    V = np.zeros((n, N))
    for j in range(N):
        idx = np.arange(j, n, N)
        V[idx, j] = c ** np.arange(idx.size, dtype=np.float64)
        V[:, j] /= np.linalg.norm(V[:, j])
    return V


def regular_incidence(graphs, verts, seed=42):
    """Oriented edge-vertex incidence of `graphs` disjoint 4-regular graphs.

    Graph g, cycle h in {0,1}: a permutation pi of 0..verts-1 from stream
    2*g+h; edge i joins tail pi[i] to head pi[(i+1) % verts]; the row holds
    +1 at the head, -1 at the tail. Rows are ordered by graph, cycle, i;
    columns of graph g are offset by g*verts. Kernel: the `graphs` component
    indicators scaled by 1/sqrt(verts).
    """
    m = graphs * 2 * verts
    n = graphs * verts
    tails = np.empty(m, np.int64)
    heads = np.empty(m, np.int64)
    r = 0
    for g in range(graphs):
        for h in range(2):
            pi = _rng(seed, 2 * g + h).permutation(verts).astype(np.int64) + g * verts
            tails[r:r + verts] = pi
            heads[r:r + verts] = np.roll(pi, -1)
            r += verts
    return _from_two_per_row(m, n, heads, np.ones(m), tails, -np.ones(m))


def component_indicators(graphs, verts):
    V = np.zeros((graphs * verts, graphs))
    for g in range(graphs):
        V[g * verts:(g + 1) * verts, g] = 1.0 / np.sqrt(verts)
    return V


def grid_incidence(grids, side, seed=42, weighted=True):
    """Weighted incidence of `grids` disjoint side x side 4-neighbour grids.

    Per grid, horizontal edges (row-major) then vertical edges; edge e from
    u to its right/lower neighbour v has +w_e at v and -w_e at u, with
    w_e = 10**(-3 U), U ~ U[0,1) from stream g (w_e = 1 if not weighted).
    Kernel: `grids` indicators scaled by 1/side.
    """
    vpg = side * side
    rows_t, rows_h, ws = [], [], []
    for g in range(grids):
        idx = np.arange(vpg, dtype=np.int64).reshape(side, side) + g * vpg
        ht, hh = idx[:, :-1].ravel(), idx[:, 1:].ravel()
        vt, vh = idx[:-1, :].ravel(), idx[1:, :].ravel()
        t = np.concatenate([ht, vt])
        h = np.concatenate([hh, vh])
        if weighted:
            w = 10.0 ** (-3.0 * _rng(seed, g).random(t.size))
        else:
            w = np.ones(t.size)
        rows_t.append(t)
        rows_h.append(h)
        ws.append(w)
    t = np.concatenate(rows_t)
    h = np.concatenate(rows_h)
    w = np.concatenate(ws)
    return _from_two_per_row(t.size, grids * vpg, h, w, t, -w)


def rg_block(m, p=32, offdiag=8, gnnz=16, seed=42):
    """A = [R | R G], m x (m+p); R has `offdiag` distinct off-diagonal N(0,1)
    entries per row at uniform columns and diagonal 1 + sum|offdiag|; G is
    m x p with `gnnz` distinct N(0,1) entries per column. Returns (A, G) with
    G dense (m x p); ker A = span [-G; I_p].
    """
    import scipy.sparse as sp
    rng = _rng(seed, 0)
    # distinct off-diagonal columns per row: sample from [0, m-1) and skip i
    cols = np.empty((m, offdiag), np.int64)
    for start in range(0, m, 1 << 18):
        stop = min(m, start + (1 << 18))
        blk = np.empty((stop - start, offdiag), np.int64)
        need = np.ones(stop - start, bool)
        while need.any():
            idx = np.nonzero(need)[0]
            c = rng.integers(0, m - 1, size=(idx.size, offdiag))
            c = c + (c >= (start + idx)[:, None])
            s = np.sort(c, axis=1)
            ok = np.all(s[:, 1:] != s[:, :-1], axis=1)
            blk[idx[ok]] = c[ok]
            need[idx[ok]] = False
        cols[start:stop] = blk
    vals = rng.standard_normal((m, offdiag))
    diag = 1.0 + np.abs(vals).sum(axis=1)
    rr = np.repeat(np.arange(m), offdiag + 1)
    cc = np.concatenate([cols, np.arange(m)[:, None]], axis=1).ravel()
    vv = np.concatenate([vals, diag[:, None]], axis=1).ravel()
    R = sp.csr_matrix((vv, (rr, cc)), shape=(m, m))
    grng = _rng(seed, 1)
    G = np.zeros((m, p))
    for j in range(p):
        rows = grng.choice(m, size=gnnz, replace=False)
        G[rows, j] = grng.standard_normal(gnnz)
    RG = R @ sp.csr_matrix(G)
    A = sp.hstack([R, RG], format="csr")
    A.sort_indices()
    A.sum_duplicates()
    return CsrMatrix(m, m + p, A.indptr.astype(np.int32), A.indices.astype(np.int32),
                     A.data.astype(np.float64)), G


def rg_kernel(G):
    m, p = G.shape
    K = np.vstack([-G, np.eye(p)])
    Q, _ = np.linalg.qr(K)
    return Q
