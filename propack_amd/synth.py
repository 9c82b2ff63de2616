"""Seeded synthetic sparse matrices for the truncated-SVD workloads.

All generators use numpy's counter-based Philox bit generator, so a seed
fixes the matrix bit-for-bit on any host; the CPU runs and the device runs
are fed the very same CSR arrays.  Matrices are returned as ``Csr`` (int32
row pointers and column indices, values in the requested dtype, columns
sorted within each row), together with their explicit transposes when asked.

Workloads (sizes as named by the benchmark configurations):

* ``random_uniform`` -- m x n with ``nnz_draws`` uniform (row, col) draws,
  duplicates summed, N(0,1) values (1850 x 712 with 8636 draws is the small
  correctness case).
* ``poisson_decay`` -- each row holds Poisson(lam) distinct uniform columns,
  value N(0,1) * (1 + j)^-p for column j (2,000,000 x 1,000,000, lam = 16,
  p = 0.6 is the large FP64 case).
* ``powerlaw_bipartite`` -- row degrees P(d) ~ d^-gamma on [1, dmax], columns
  from a Zipf(s) law over a seeded permutation, values 1.0.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Csr:
    m: int
    n: int
    rowptr: np.ndarray  # int32, length m + 1
    col: np.ndarray     # int32, length nnz
    val: np.ndarray     # float64, float32, complex128 or complex64, length nnz

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def astype(self, dtype) -> "Csr":
        return Csr(self.m, self.n, self.rowptr, self.col,
                   np.ascontiguousarray(self.val, dtype=dtype))

    def transpose(self) -> "Csr":
        """Explicit CSR of A^T (stable counting sort by column)."""
        rows = np.repeat(np.arange(self.m, dtype=np.int32),
                         np.diff(self.rowptr).astype(np.int64))
        order = np.argsort(self.col, kind="stable")
        counts = np.bincount(self.col, minlength=self.n)
        rp = np.zeros(self.n + 1, dtype=np.int64)
        np.cumsum(counts, out=rp[1:])
        return Csr(self.n, self.m, rp.astype(np.int32),
                   rows[order].astype(np.int32), self.val[order].copy())

    @property
    def is_complex(self) -> bool:
        return np.iscomplexobj(self.val)

    def _wide(self):
        return np.complex128 if self.is_complex else np.float64

    def todense(self) -> np.ndarray:
        a = np.zeros((self.m, self.n), dtype=self._wide())
        rows = np.repeat(np.arange(self.m), np.diff(self.rowptr))
        np.add.at(a, (rows, self.col), self.val.astype(self._wide()))
        return a

    def matvec(self, x: np.ndarray) -> np.ndarray:
        """y = A x in float64 / complex128 (host helper for checks, not a
        product path)."""
        rows = np.repeat(np.arange(self.m), np.diff(self.rowptr))
        wide = np.result_type(self._wide(), np.asarray(x).dtype)
        y = np.zeros(self.m, dtype=wide)
        np.add.at(y, rows, self.val.astype(wide) * x[self.col])
        return y


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def from_coo(m: int, n: int, rows: np.ndarray, cols: np.ndarray,
             vals: np.ndarray, sum_duplicates: bool = True) -> Csr:
    """COO -> CSR with columns sorted inside each row, duplicates summed."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals)
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    key = key[order]
    vals = vals[order]
    if sum_duplicates and key.size:
        first = np.ones(key.size, dtype=bool)
        first[1:] = key[1:] != key[:-1]
        seg = np.cumsum(first) - 1
        wide = np.complex128 if np.iscomplexobj(vals) else np.float64
        summed = np.zeros(int(first.sum()), dtype=wide)
        np.add.at(summed, seg, vals.astype(wide))
        key = key[first]
        vals = summed.astype(vals.dtype)
    r = (key // n).astype(np.int64)
    c = (key % n).astype(np.int32)
    counts = np.bincount(r, minlength=m)
    rp = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(counts, out=rp[1:])
    if rp[-1] >= 2**31:
        raise ValueError("nnz exceeds int32 row pointers")
    return Csr(m, n, rp.astype(np.int32), c, np.ascontiguousarray(vals))


def random_uniform(m: int = 1850, n: int = 712, nnz_draws: int = 8636,
                   seed: int = 1850, dtype=np.float64) -> Csr:
    g = _rng(seed)
    rows = g.integers(0, m, nnz_draws)
    cols = g.integers(0, n, nnz_draws)
    vals = g.standard_normal(nnz_draws)
    return from_coo(m, n, rows, cols, vals).astype(dtype)


def complex_uniform(m: int = 1850, n: int = 712, nnz_draws: int = 8636,
                    seed: int = 1850, dtype=np.complex128) -> Csr:
    """random_uniform's sparsity pattern with complex N(0,1) + i N(0,1)
    values (the imaginary parts drawn after the real ones)."""
    g = _rng(seed)
    rows = g.integers(0, m, nnz_draws)
    cols = g.integers(0, n, nnz_draws)
    vals = g.standard_normal(nnz_draws) + 1j * g.standard_normal(nnz_draws)
    return from_coo(m, n, rows, cols, vals).astype(dtype)


def complex_poisson_decay(m: int = 1_000_000, n: int = 500_000, lam: float = 16.0,
                          p: float = 0.6, seed: int = 2025, dtype=np.complex128) -> Csr:
    """poisson_decay's pattern and decay with each value multiplied by a
    seeded unit phase exp(i theta), theta uniform on [0, 2 pi)."""
    A = poisson_decay(m, n, lam, p, seed=seed, dtype=np.float64)
    theta = _rng(seed + 1).uniform(0.0, 2.0 * np.pi, A.nnz)
    return Csr(A.m, A.n, A.rowptr, A.col, (A.val * np.exp(1j * theta)).astype(dtype))


def poisson_decay(m: int = 2_000_000, n: int = 1_000_000, lam: float = 16.0,
                  p: float = 0.6, seed: int = 2024, dtype=np.float64) -> Csr:
    g = _rng(seed)
    counts = np.minimum(g.poisson(lam, m), n).astype(np.int64)
    rp = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(counts, out=rp[1:])
    nnz = int(rp[-1])
    if nnz >= 2**31:
        raise ValueError("nnz exceeds int32 row pointers")
    rows = np.repeat(np.arange(m, dtype=np.int64), counts)
    cols = g.integers(0, n, nnz)
    # Distinct columns per row: redraw the later copy of each duplicate.
    for _ in range(64):
        key = rows * n + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(nnz, dtype=bool)
        dup[1:] = ks[1:] == ks[:-1]
        if not dup.any():
            break
        idx = order[dup]
        cols[idx] = g.integers(0, n, idx.size)
    else:
        raise RuntimeError("could not make columns distinct")
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    cols = cols[order].astype(np.int32)
    vals = g.standard_normal(nnz) * np.power(1.0 + cols.astype(np.float64), -p)
    return Csr(m, n, rp.astype(np.int32), cols, vals.astype(dtype))


def _zipf_distinct(g: np.random.Generator, deg: np.ndarray, rcdf: np.ndarray) -> np.ndarray:
    """0-based Zipf ranks, deg[i] distinct ones for row i, in row order.

    Successive sampling: i.i.d. draws from the law whose CDF is rcdf, a draw
    kept when its rank is new to its row, until the row holds deg[i] ranks.
    The draws of a round are taken together; a later copy of a rank inside
    a row is redrawn in the next round.  A redraw is conditioned on a rank
    above the row's leading run of taken ranks 0..K-1: a draw inside the
    run would be rejected anyway, so the law of the set is unchanged, and
    rows of 10^5 entries (whose taken ranks cover a long prefix) finish in
    a few dozen rounds instead of ~10^9 draws."""
    n = rcdf.size
    m = deg.size
    nnz = int(deg.sum())
    rp = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(deg, out=rp[1:])
    ranks = np.minimum(np.searchsorted(rcdf, g.random(nnz), side="right"), n - 1)
    # entries of the rows still holding duplicates, as (flat index, row)
    idx = np.arange(nnz, dtype=np.int64)
    rows = np.repeat(np.arange(m, dtype=np.int64), deg)
    multi = deg[rows] > 1
    idx, rows = idx[multi], rows[multi]
    for _ in range(100_000):
        if idx.size == 0:
            return ranks
        key = rows * n + ranks[idx]
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(ks.size, dtype=bool)
        dup[1:] = ks[1:] == ks[:-1]
        if not dup.any():
            return ranks
        drows = np.unique(rows[order[dup]])
        # keep only the rows with duplicates (sorted order is by row)
        keep = np.isin(rows[order], drows)
        order, ks, dup = order[keep], ks[keep], dup[keep]
        srows = rows[order]
        sr = ks - srows * n                     # ranks, sorted within each row
        # leading run K of each row: distinct ranks 0..K-1 at its first K
        # distinct positions
        first = np.ones(srows.size, dtype=bool)
        first[1:] = srows[1:] != srows[:-1]
        start = np.flatnonzero(first)
        uniq = ~dup
        cum = np.cumsum(uniq) - 1               # distinct position so far
        pos = cum - cum[start][np.cumsum(first) - 1]
        inrun = uniq & (sr == pos)
        brk = np.where(uniq & ~inrun, pos, np.iinfo(np.int64).max)
        kk = np.minimum.reduceat(brk, start)
        ndist = np.add.reduceat(uniq.astype(np.int64), start)
        K = np.minimum(kk, ndist)               # per row of drows
        ridx = np.searchsorted(drows, srows[dup])
        lo = np.where(K[ridx] > 0, rcdf[np.maximum(K[ridx] - 1, 0)], 0.0)
        u = lo + (1.0 - lo) * g.random(lo.size)
        ranks[idx[order[dup]]] = np.minimum(np.searchsorted(rcdf, u, side="right"), n - 1)
        idx, rows = idx[order], srows
    raise RuntimeError("could not make columns distinct")


def powerlaw_bipartite(m: int = 8_000_000, n: int = 4_000_000,
                       gamma: float = 2.1, dmax: int = 100_000,
                       zipf_s: float = 1.8, seed: int = 8000,
                       dtype=np.float32) -> Csr:
    """Row degrees d ~ d^-gamma on [1, dmax]; each row holds d distinct
    columns drawn from a Zipf(zipf_s) law over a seeded permutation of the
    column ids (successive sampling, _zipf_distinct); every value is 1.0."""
    g = _rng(seed)
    # Row degrees by inverse-CDF sampling of the discrete law d^-gamma.
    d = np.arange(1, dmax + 1, dtype=np.float64)
    cdf = np.cumsum(d ** -gamma)
    cdf /= cdf[-1]
    deg = np.searchsorted(cdf, g.random(m), side="right") + 1
    deg = np.minimum(deg, n).astype(np.int64)
    # Column ids: Zipf(s) rank over a seeded permutation of the columns.
    perm = g.permutation(n).astype(np.int64)
    rcdf = np.cumsum(np.arange(1, n + 1, dtype=np.float64) ** -zipf_s)
    rcdf /= rcdf[-1]
    ranks = _zipf_distinct(g, deg, rcdf)
    rp = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(deg, out=rp[1:])
    if rp[-1] >= 2**31:
        raise ValueError("nnz exceeds int32 row pointers")
    rows = np.repeat(np.arange(m, dtype=np.int64), deg)
    cols = perm[ranks]
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    return Csr(m, n, rp.astype(np.int32), cols[order].astype(np.int32),
               np.ones(int(rp[-1]), dtype=dtype))


def batch_params(count: int = 256, m_lo: int = 5_000, m_hi: int = 500_000,
                 seed: int = 4096) -> list:
    """Shapes of the batched workload, one dict per problem: m log-uniform
    in [m_lo, m_hi], n = round(m * r) with r ~ U[0.25, 1], Poisson(lam)
    nonzeros per row with lam ~ U[4, 48], column decay p ~ U[0.3, 1.0], and
    the seed of its poisson_decay matrix."""
    g = _rng(seed)
    m = np.exp(g.uniform(np.log(m_lo), np.log(m_hi), count)).round().astype(np.int64)
    r = g.uniform(0.25, 1.0, count)
    n = np.maximum(1, np.round(m * r)).astype(np.int64)
    lam = g.uniform(4.0, 48.0, count)
    p = g.uniform(0.3, 1.0, count)
    return [dict(m=int(m[i]), n=int(n[i]), lam=float(lam[i]), p=float(p[i]),
                 seed=int(seed + 1 + i)) for i in range(count)]


def batch_problem(prm: dict, dtype=np.float64) -> Csr:
    """The matrix of one batched problem (see batch_params)."""
    return poisson_decay(prm["m"], prm["n"], prm["lam"], prm["p"], seed=prm["seed"], dtype=dtype)


def start_vector(m: int, seed: int, dtype=np.float64) -> np.ndarray:
    """Seeded N(0,1) start vector of length m (complex dtypes: N(0,1) real
    parts, then N(0,1) imaginary parts)."""
    g = _rng(seed)
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        re = g.standard_normal(m)
        return (re + 1j * g.standard_normal(m)).astype(dtype)
    return g.standard_normal(m).astype(dtype)
