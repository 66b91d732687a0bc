"""The bound of ||Lap_h^-1||inf the GPU comparisons use (tests/bounds.py),
checked against the exact inverse of the assembled ghost-node Laplacian on
small grids for every mix of face types."""
import itertools

import numpy as np
import pytest

import bounds


def dense_laplacian(n, h, bc):
    """7-point Lap_h on the unknowns of a grid (Dirichlet boundary nodes
    eliminated with value 0, Neumann boundary nodes kept, their ghost = the
    mirror point)."""
    lo = [bounds.W, bounds.S, bounds.B]
    hi = [bounds.E, bounds.N, bounds.T]
    rng = [range(int(bc[lo[a]] == "D"), n[a] - int(bc[hi[a]] == "D")) for a in range(3)]
    pts = list(itertools.product(*rng))
    index = {p: m for m, p in enumerate(pts)}
    A = np.zeros((len(pts), len(pts)))
    for p, m in index.items():
        for a in range(3):
            for d in (-1, 1):
                q = list(p)
                q[a] += d
                if not 0 <= q[a] < n[a]:
                    q[a] -= 2 * d  # ghost = mirror (homogeneous Neumann)
                A[m, m] -= 1.0 / h ** 2
                q = tuple(q)
                if q in index:
                    A[m, index[q]] += 1.0 / h ** 2
    return A


BCS = ["DDDDDD", "NNNNDD", "NDNDND", "DNNDNN", "NNNNND", "DDNNNN"]


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("n", [(5, 5, 5), (7, 5, 4), (4, 6, 9)])
def test_linv_bound_holds(n, bc):
    h = 1.0 / (n[0] - 1)
    A = dense_laplacian(n, h, bc)
    norm = np.abs(np.linalg.inv(A)).sum(axis=1).max()
    bound = bounds.linv_bound(n, h, bc)
    assert norm <= bound * (1 + 1e-12), (norm, bound)
    # and it is not loose by more than an order of magnitude on these grids
    assert norm >= 0.05 * bound


def test_all_neumann_is_rejected():
    with pytest.raises(ValueError):
        bounds.linv_bound((5, 5, 5), 0.25, "NNNNNN")
