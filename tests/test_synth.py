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
