"""Argument checks of the svds front end (scipy's ``_iv`` for the propack
solver, scipy/sparse/linalg/_eigen/_svds.py): they run before any device
work, so they need no GPU."""
import numpy as np
import pytest

from propack_amd.svdp import svds


@pytest.mark.parametrize("kwargs, match", [
    ({"k": 0}, "`k` must be an integer"),
    ({"k": 11}, "`k` must be an integer"),
    ({"k": 2.5}, "`k` must be an integer"),
    ({"tol": -1.0}, "`tol` must be a non-negative"),
    ({"tol": np.inf}, "`tol` must be a non-negative"),
    ({"which": "LA"}, "`which` must be in"),
    ({"v0": np.ones(7)}, "`v0` must have shape"),
    ({"v0": np.ones(10, dtype=np.int64)}, "`v0` must be of floating"),
    ({"maxiter": 0}, "`maxiter` must be a positive integer"),
    ({"maxiter": 1.5}, "`maxiter` must be a positive integer"),
    ({"return_singular_vectors": "x"}, "`return_singular_vectors` must be in"),
    ({"solver": "arpack"}, "solver must be 'propack'"),
])
def test_svds_argument_errors(kwargs, match):
    A = np.arange(1.0, 201.0).reshape(10, 20)
    args = {"k": 3, **kwargs}
    with pytest.raises(ValueError, match=match):
        svds(A, **args)


def test_svds_rejects_empty_and_non_numeric():
    with pytest.raises(ValueError, match="must not be empty"):
        svds(np.zeros((0, 4)), k=1)
    with pytest.raises(ValueError, match="numeric"):
        svds(np.array([["a", "b"], ["c", "d"]]), k=1)
