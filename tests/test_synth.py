"""Seeded synthetic workloads: shapes, determinism and distribution facts."""
import numpy as np

from propack_amd import synth


def test_random_uniform_small_case():
    A = synth.random_uniform()
    assert (A.m, A.n) == (1850, 712)
    assert 8500 < A.nnz <= 8636   # 8636 draws, duplicates summed
    B = synth.random_uniform()
    assert np.array_equal(A.col, B.col) and np.array_equal(A.val, B.val)
    # columns sorted within rows
    for i in range(0, A.m, 97):
        c = A.col[A.rowptr[i]:A.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0)


def test_transpose_roundtrip():
    A = synth.random_uniform(300, 200, 3000, seed=7)
    T = A.transpose()
    assert np.array_equal(T.transpose().todense(), A.todense())
    assert np.allclose(T.todense(), A.todense().T)


def test_poisson_decay_small():
    A = synth.poisson_decay(20000, 10000, 16.0, 0.6, seed=3)
    deg = np.diff(A.rowptr)
    assert abs(deg.mean() - 16.0) < 0.2
    for i in range(0, A.m, 501):
        c = A.col[A.rowptr[i]:A.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0)          # distinct, sorted
    # the column scaling (1 + j)^-0.6 shrinks late columns
    T = A.transpose()
    early = np.abs(T.val[:T.rowptr[100]]).mean()
    late = np.abs(T.val[T.rowptr[-101]:]).mean()
    assert late < early * (100 / 9900) ** 0.6 * 3


def test_start_vector_seeded():
    assert np.array_equal(synth.start_vector(10, 1), synth.start_vector(10, 1))
    assert not np.array_equal(synth.start_vector(10, 1), synth.start_vector(10, 2))


def _naive_distinct(g, deg, rcdf):
    """Successive sampling by plain rejection, one row at a time."""
    out = []
    for d in deg:
        taken = []
        while len(taken) < d:
            r = int(min(np.searchsorted(rcdf, g.random(), side="right"), rcdf.size - 1))
            if r not in taken:
                taken.append(r)
        out.append(sorted(taken))
    return out


def test_zipf_distinct_law_matches_rejection():
    # inclusion frequency of every rank, generator vs plain rejection, on
    # rows whose degree forces long runs of taken ranks (the conditioned
    # redraws); a wrong run length K biases the low ranks far beyond noise
    n, rows = 24, 6000
    rcdf = np.cumsum(np.arange(1, n + 1, dtype=np.float64) ** -1.8)
    rcdf /= rcdf[-1]
    deg = np.fromiter(([1, 3, 8, 15][i % 4] for i in range(rows)), dtype=np.int64)
    r = synth._zipf_distinct(synth._rng(5), deg, rcdf)
    rp = np.concatenate([[0], np.cumsum(deg)])
    for d in (3, 8, 15):
        sel = [i for i in range(rows) if deg[i] == d]
        fast = np.zeros(n)
        for i in sel:
            seg = r[rp[i]:rp[i + 1]]
            assert len(set(seg.tolist())) == d
            fast[seg] += 1
        slow = np.zeros(n)
        for s in _naive_distinct(synth._rng(6), deg[sel], rcdf):
            slow[s] += 1
        p = (fast + slow) / (2 * len(sel))
        sd = np.sqrt(np.maximum(p * (1 - p), 1e-12) * 2 / len(sel))
        z = np.abs(fast - slow) / len(sel) / sd
        assert z.max() < 5.0, (d, z.max())


def test_powerlaw_distinct_unit_values():
    A = synth.powerlaw_bipartite(60000, 30000, seed=11)
    deg = np.diff(A.rowptr)
    assert np.all(A.val == 1.0)
    for i in np.argsort(deg)[-20:]:
        c = A.col[A.rowptr[i]:A.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0)          # distinct, sorted
    # degree law d^-2.1 on [1, 1e5] (clipped at n): mean about 4.6
    assert 3.5 < deg.mean() < 6.0
    assert A.nnz == int(deg.sum())
