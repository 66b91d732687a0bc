"""Second-order convergence on the GPU for every solve path: the max-abs
error against the analytic solution falls by 3.5-4.5x per halving of h
(reference README.md:142-146, convergence.bsub:23-36; the reference's own
MI250X runs show the same ratio, results/MI250X/dirichlet_1.txt:2-14).

Paths: the FP64 consistent-mode F-cycle; FP32 F-cycles inside FP64
refinement (hi + lo iterate, FP64 coarsest level), with and without FP16
coarse levels; the batched solver; the serving loop.  Problems: u = cos(x z^2)
sin(y^3) with Dirichlet faces (the mixed configuration's problem) and
u = x^3 y^2 z with Neumann x/y faces and Dirichlet z faces (the headline
problem), on n x n x (n+1)/2 grids with h = 1/(n-1); for the batched and
served paths, zero-Dirichlet cubes with f a sum of Fourier modes, whose
solution is the same sum divided by -pi^2 (a^2 + b^2 + c^2).
"""
import zlib

import numpy as np
import pytest

import mgp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def exact(problem, dims, h):
    X, Y, Z = np.meshgrid(*[np.arange(n) * h for n in dims], indexing="ij")
    if problem == 3:
        return np.cos(X * Z * Z) * np.sin(Y ** 3)
    return X ** 3 * Y ** 2 * Z


def check_ratios(errs, label):
    ratios = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
    say("  %s errors %s ratios %s" % (label, ["%.3e" % e for e in errs], ["%.3f" % r for r in ratios]))
    assert all(3.5 < r < 4.5 for r in ratios), (errs, ratios)


PATHS = [("fp64", 3, "DDDDDD", 0), ("fp64", 2, "NNNNDD", 0), ("mixed", 3, "DDDDDD", 0), ("mixed", 2, "NNNNDD", 0),
         ("mixed", 3, "DDDDDD", 2)]


@pytest.mark.parametrize("path,problem,bc,half_from", PATHS, ids=["%s-p%d-%s-h%d" % p for p in PATHS])
def test_single_solver_second_order(path, problem, bc, half_from):
    errs = []
    sizes = (33, 65, 129, 257) if path == "mixed" and problem == 3 else (33, 65, 129)
    for n in sizes:
        dims = (n, n, (n + 1) // 2)
        h = 1.0 / (n - 1)
        levels = 4
        say("case convergence", path, problem, bc, half_from, dims)
        dtype = mgp.F64 if path == "fp64" else mgp.MIXED
        s = mgp.Solver(dims, levels, h, (0, 0, 0), bc, mgp.CONSISTENT, mgp.GAUSS_SEIDEL, mgp.FULL_WEIGHTING, problem,
                       dtype=dtype, maxiter=80, miniter=1, tolerance=1e-10, coarse_sweeps=16, half_from=half_from)
        s.to_device()
        it, rel, _ = s.solve("Fcycle", 2, 1.0)
        assert rel <= 1e-10, (n, it, rel)
        s.to_host()
        u = np.array(s.array("u"), copy=True)
        L = s.layout()
        hs = list(L.halo)
        u = u.reshape(dims[0] + hs[2] + hs[3], dims[1] + hs[0] + hs[1], dims[2] + hs[4] + hs[5])
        u = u[hs[3]:hs[3] + dims[0], hs[1]:hs[1] + dims[1], hs[5]:hs[5] + dims[2]]
        errs.append(np.abs(u - exact(problem, dims, h)).max())
        s.close()
    check_ratios(errs, "%s p%d %s h%d" % (path, problem, bc, half_from))


def modes_for(seed):
    rng = np.random.default_rng(seed)
    return [(int(a), int(b), int(c), float(rng.standard_normal())) for a, b, c in rng.integers(1, 4, (4, 3))]


def fourier(n, modes, solution=False):
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    out = np.zeros((n, n, n))
    for a, b, c, A in modes:
        w = A / (-np.pi ** 2 * (a * a + b * b + c * c)) if solution else A
        out += w * np.sin(a * np.pi * X) * np.sin(b * np.pi * Y) * np.sin(c * np.pi * Z)
    return out


def test_batched_second_order():
    modes = modes_for(zlib.crc32(b"conv-batch"))
    sizes = (17, 33, 65, 129)
    dims = [(n, n, n) for n in sizes]
    b = mgp.Batch(dims, relaxation=mgp.JACOBI, nsmooth=3, omega=6.0 / 7.0, maxiter=60, tolerance=1e-10)
    for p, n in enumerate(sizes):
        b.set_rhs(p, fourier(n, modes))
    b.solve()
    errs = []
    for p, n in enumerate(sizes):
        it, rel = b.stats(p)
        assert rel <= 1e-10
        errs.append(np.abs(b.solution(p) - fourier(n, modes, solution=True)).max())
    b.close()
    check_ratios(errs, "batched")


def test_served_second_order():
    modes = modes_for(zlib.crc32(b"conv-serve"))
    sizes = (17, 33, 65, 129)
    s = mgp.Server(max_active=4, tolerance=1e-8)
    ids = [s.submit((n, n, n), "DDDDDD", modes) for n in sizes]
    errs = []
    for rid, n in zip(ids, sizes):
        res, u = s.wait(rid)
        assert res.status == 0 and res.rel_res <= 1e-8
        errs.append(np.abs(u - fourier(n, modes, solution=True)).max())
    s.close()
    check_ratios(errs, "served")
