"""Batched solves (mgp_batch_*) on the GPU: every problem of a batch ends with
the same bits (solution, iteration count, relative residual) as the same
problem solved alone by the single-problem mixed-precision solver with the
batch's arithmetic (plain FP64 iterate, FP32 coarsest level:
MGP_SOLVER_PLAIN_ITERATE | MGP_SOLVER_COARSE_FP32).  The batch against the
oracle's FP64 solves: tests/test_gpu_refinement.py."""
import ctypes
import zlib

import numpy as np
import pytest

import mgp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def max_levels(d):
    """Levels down to a 3-point coarsest axis with exact 2:1 nesting."""
    L = 1
    while all(((n - 1) >> (L - 1)) % 2 == 0 and ((n - 1) >> (L - 1)) // 2 >= 2 for n in d):
        L += 1
    return L


def fourier_rhs(d, rng):
    """f = sum of 4 modes A sin(a pi x) sin(b pi y) sin(c pi z), a, b, c in 1..8,
    A ~ N(0, 1), on the grid x = i h (h = 1/(nx-1)) over the whole shape."""
    h = 1.0 / (d[0] - 1)
    x = np.arange(d[0])[:, None, None] * h
    y = np.arange(d[1])[None, :, None] * h
    z = np.arange(d[2])[None, None, :] * h
    f = np.zeros(d)
    for _ in range(4):
        a, b, c = rng.integers(1, 9, size=3)
        f += rng.standard_normal() * np.sin(a * np.pi * x) * np.sin(b * np.pi * y) * np.sin(c * np.pi * z)
    return f


def single_solve(d, f, relaxation, nsmooth, omega, coarse_sweeps, maxiter, tol):
    s = mgp.Solver(d, max_levels(d), 1.0 / (d[0] - 1), (0.0, 0.0, 0.0), "DDDDDD", mgp.CONSISTENT, relaxation,
                   mgp.FULL_WEIGHTING, problem=-1, dtype=mgp.MIXED, maxiter=maxiter, miniter=1, tolerance=tol,
                   coarse_sweeps=coarse_sweeps, flags=mgp.QUIET | mgp.PLAIN_ITERATE | mgp.COARSE_FP32)
    s.array("f")[:] = f.ravel()
    s.to_device()
    it, rel, _ = s.solve("Vcycle", nsmooth, omega)
    s.to_host()
    u = s.array("u").copy().reshape(d)
    s.close()
    return it, rel, u


SHAPES = [(9, 9, 9), (17, 17, 17), (33, 17, 9), (9, 17, 33), (33, 33, 33), (17, 33, 65)]


@pytest.mark.parametrize("relaxation,nsmooth,omega,cs", [(mgp.JACOBI, 3, 6.0 / 7.0, 0), (mgp.GAUSS_SEIDEL, 2, 1.0, 8)],
                         ids=["jacobi", "gs"])
@pytest.mark.parametrize("graphs", [True, False], ids=["graphs", "nographs"])
def test_batch_equals_single_solves_bitwise(relaxation, nsmooth, omega, cs, graphs):
    say("case batch", relaxation, graphs)
    rng = np.random.default_rng(zlib.crc32(b"batch"))
    fs = [fourier_rhs(d, rng) for d in SHAPES]
    tol = 1e-10
    b = mgp.Batch(SHAPES, relaxation=relaxation, nsmooth=nsmooth, omega=omega, coarse_sweeps=cs, maxiter=40,
                  tolerance=tol, flags=0 if graphs else mgp.BATCH_NO_GRAPHS)
    for p, f in enumerate(fs):
        b.set_rhs(p, f)
    b.solve()
    its = set()
    for p, (d, f) in enumerate(zip(SHAPES, fs)):
        it, rel = b.stats(p)
        u = b.solution(p)
        sit, srel, su = single_solve(d, f, relaxation, nsmooth, omega, cs, 40, tol)
        assert (it, rel) == (sit, srel), (d, it, rel, sit, srel)
        assert rel <= tol
        np.testing.assert_array_equal(u, su)
        its.add(it)
    # the batch mixes problems that stop after different iteration counts
    say("iterations", sorted(its))
    # a second solve from u = 0 repeats the first exactly
    u0 = [b.solution(p) for p in range(len(SHAPES))]
    st0 = [b.stats(p) for p in range(len(SHAPES))]
    b.solve()
    for p in range(len(SHAPES)):
        assert b.stats(p) == st0[p]
        np.testing.assert_array_equal(b.solution(p), u0[p])
    b.close()


def test_batch_tiles_come_from_the_candidates():
    b = mgp.Batch([(33, 33, 33), (65, 33, 17)], maxiter=1)
    cands = {(1, 4), (2, 4), (4, 4), (8, 4), (16, 4), (1, 8), (2, 8), (4, 8), (8, 8), (1, 16), (4, 16)}
    for p, d in enumerate([(33, 33, 33), (65, 33, 17)]):
        for level in range(max_levels(d)):
            assert b.tile(p, level) in cands
    with pytest.raises(mgp.MgpError):
        b.tile(0, 99)
    b.close()


def test_batch_rejects_bad_shapes():
    with pytest.raises(mgp.MgpError):
        mgp.Batch([(2, 9, 9)])
    with pytest.raises(mgp.MgpError):
        mgp.Batch([(9, 9, 9)], levels=7)
