"""The CPU oracle, pinned before it is trusted (CPU only, no GPU).

Reference mode is pinned bit for bit against golden vectors recorded from the
reference itself (tests/golden/, made by tests/golden/make_goldens.py from the
reference compiled unmodified).  Consistent mode -- which the reference does
not have -- is pinned by what the mathematics fixes: exactness of the 7-point
stencil on cubics, eigen-damping of Jacobi, the ghost-node Neumann identity, a
dense direct solve of the discrete system on small grids, and second-order
convergence to the analytic solution.
"""
import ctypes

import numpy as np
import pytest

import goldens
import oracle_lib as O


# ----------------------------------------------------------- golden parity

SOLVE_CASES = [n for n in goldens.names("g") if goldens.load(n)["params"]["dtype"] == "f64"]


@pytest.mark.parametrize("name", SOLVE_CASES)
def test_solve_matches_reference_bitwise(name):
    g = goldens.load(name)
    p = g["params"]
    kw = goldens.solve_kwargs(p)
    rk = goldens.run_kwargs(p)
    s = O.Solver(kw["dims"], kw["levels"], kw["h"], kw["origin"], kw["bc"], O.REFERENCE, kw["relaxation"],
                 kw["restriction"], kw["problem"])
    s.to_device()
    # the recording called solve() once per cycle with maxiter = 1 (no stop rule)
    rr = []
    for _ in range(rk["ncycles"]):
        it, fin, _ = s.solve(O.cycle_id(rk["cycle"]), rk["nsmooth"], rk["omega"], maxiter=1, miniter=1, tol=0.0)
        assert it == 1
        rr.append(fin)
    np.testing.assert_array_equal(np.array(rr), g["relres"])
    L = s.layout(0)
    shp = O.padded_shape(L)
    assert goldens.sha(s.array("f").reshape(shp)) == str(g["f_sha"])
    assert goldens.sha(s.array("u").reshape(shp)) == str(g["u_sha"])
    assert goldens.sha(s.array("r").reshape(shp)) == str(g["r_sha"])


@pytest.mark.parametrize("name", goldens.names("stop_"))
def test_stop_rule_matches_reference(name):
    g = goldens.load(name)
    p = g["params"]
    kw = goldens.solve_kwargs(p)
    s = O.Solver(kw["dims"], kw["levels"], kw["h"], kw["origin"], kw["bc"], O.REFERENCE, kw["relaxation"],
                 kw["restriction"], kw["problem"])
    s.to_device()
    it, fin, _ = s.solve(O.cycle_id(p["cycle"]), int(p["nsmooth"]), float(p["omega"]), int(p["maxiter"]),
                         int(p["miniter"]), float(p["tol"]))
    assert it == int(g["iterations"])
    assert fin == float(g["rel_res"])
    assert goldens.sha(s.array("u")) == str(g["u_sha"])


def _level_solver(p, relaxation):
    dims = (int(p["nx"]), int(p["ny"]), int(p["nz"]))
    return O.Solver(dims, 1, 1.0 / (dims[0] - 1.0), (0, 0, 0), p["bc"], O.REFERENCE, relaxation,
                    O.INJECTION, -1)


def _load_level(s, g, with_prev):
    s.array("u")[:] = g["in_u"].ravel()
    if with_prev:
        s.array("p")[:] = g["in_uprev"].ravel()
    s.array("f")[:] = g["in_f"].ravel()
    s.array("r")[:] = g["in_r"].ravel()
    for q in range(6):
        s.face(q)[:] = g["in_g" + O.LOC_NAMES[q]].ravel()


@pytest.mark.parametrize("name", goldens.names("op_residual") + goldens.names("op_relax") +
                         goldens.names("op_gs_kernel"))
def test_level_operator_matches_reference_bitwise(name):
    g = goldens.load(name)
    p = g["params"]
    jac = p["op"] == "relax_jacobi"
    s = _level_solver(p, O.JACOBI if jac else O.GAUSS_SEIDEL)
    _load_level(s, g, jac)
    if p["op"] == "residual":
        s.residual(0)
    elif p["op"] in ("relax_gs", "relax_jacobi"):
        s.relax(0, float(p["omega"]), int(p["nsmooth"]))
    else:
        L = s.layout(0)
        box = [int(s.face_type(O.W) == O.DIRICHLET), L.n[0] - int(s.face_type(O.E) == O.DIRICHLET),
               int(s.face_type(O.S) == O.DIRICHLET), L.n[1] - int(s.face_type(O.N) == O.DIRICHLET),
               int(s.face_type(O.B) == O.DIRICHLET), L.n[2] - int(s.face_type(O.T) == O.DIRICHLET)]
        u = s.array("u")
        f = s.array("f")
        for hs in range(int(p["halfsweeps"])):
            O.lib().og_gs_half(O.REFERENCE, ctypes.byref(L), O.ptr(u), O.ptr(f), O.box_arr(box),
                               float(p["omega"]), s.level_h(0), hs % 2)
    np.testing.assert_array_equal(s.array("u"), g["out_u"].ravel())
    np.testing.assert_array_equal(s.array("r"), g["out_r"].ravel())
    if jac:
        np.testing.assert_array_equal(s.array("p"), g["out_uprev"].ravel())


def _lay(p, pre):
    n = (int(p[pre + "_nx"]), int(p[pre + "_ny"]), int(p[pre + "_nz"]))
    h = [int(p[pre + "_h" + c]) for c in O.LOC_NAMES]
    return O.layout(n, h)


@pytest.mark.parametrize("name", goldens.names("op_restrict") + goldens.names("op_prolong"))
def test_transfer_operator_matches_reference_bitwise(name):
    g = goldens.load(name)
    p = g["params"]
    Li, Lo = _lay(p, "in"), _lay(p, "out")
    a = np.ascontiguousarray(g["in_a"].ravel())
    b = np.ascontiguousarray(g["in_b"].ravel())
    fn = {"restrict_inj": O.lib().og_restrict_inject, "restrict_fw": O.lib().og_restrict_fw,
          "prolong": O.lib().og_prolong_trilinear}[p["op"]]
    fn(ctypes.byref(Li), O.ptr(a), ctypes.byref(Lo), O.ptr(b))
    np.testing.assert_array_equal(b, g["out_b"].ravel())


@pytest.mark.parametrize("name", goldens.names("op_bc_restrict"))
def test_boundary_restriction_matches_reference_bitwise(name):
    g = goldens.load(name)
    p = g["params"]
    dims = (int(p["nx"]), int(p["ny"]), int(p["nz"]))
    halo = [1 if c == "N" else 0 for c in p["bc"]]
    Lf = O.layout(dims, halo)
    hf = 1.0 / (dims[0] - 1.0)
    u = np.ascontiguousarray(g["in_u"].ravel())
    eff = list(p["bc"])
    eff[O.N] = eff[O.S] = eff[O.B] = "D"  # domain.h:139-145, 191-192
    restr = O.FULL_WEIGHTING if p["restriction"] == "fw" else O.INJECTION
    for q in range(6):
        gc = np.ascontiguousarray(g["in_cg" + O.LOC_NAMES[q]].ravel())
        if eff[q] == "D":
            gc[:] = 0.0
        else:
            gf = np.ascontiguousarray(g["in_g" + O.LOC_NAMES[q]].ravel())
            O.lib().og_bc_restrict_neumann(restr, ctypes.byref(Lf), O.ptr(u), q, O.ptr(gf), O.ptr(gc), hf)
        np.testing.assert_array_equal(gc, g["out_cg" + O.LOC_NAMES[q]].ravel())


# ------------------------------------------------ mathematical pins (both modes)

def _coords(s, level=0):
    L = s.layout(level)
    h = s.level_h(level)
    o = s.params.origin
    x = o[0] + np.arange(L.n[0]) * h
    y = o[1] + np.arange(L.n[1]) * h
    z = o[2] + np.arange(L.n[2]) * h
    return np.meshgrid(x, y, z, indexing="ij")


def _fill_interior(s, which, vals, level=0):
    L = s.layout(level)
    O.interior(L, s.array(which, level))[...] = vals


@pytest.mark.parametrize("mode", [O.REFERENCE, O.CONSISTENT])
def test_residual_vanishes_on_cubic(mode):
    """The 7-point Laplacian is exact for x^3 y^2 z: r = f - L_h u is rounding only."""
    n = 17
    s = O.Solver((n, n, n), 1, 1.0 / (n - 1), (0.3, 0.1, 0.2), "DDDDDD", mode, O.GAUSS_SEIDEL, O.INJECTION, 2)
    X, Y, Z = _coords(s)
    _fill_interior(s, "u", X ** 3 * Y ** 2 * Z)
    s.residual(0)
    r = O.interior(s.layout(), s.array("r"))[1:-1, 1:-1, 1:-1]
    fmax = np.abs(s.array("f")).max()
    assert np.abs(r).max() / fmax < 1e-10
    # a non-polynomial has an O(h^2) truncation error, so the same check fails
    # (guards against a residual that ignores u)
    _fill_interior(s, "u", np.sin(3 * X) * Y ** 2 * Z)
    s.residual(0)
    r = O.interior(s.layout(), s.array("r"))[1:-1, 1:-1, 1:-1]
    assert np.abs(r).max() / fmax > 1e-4


@pytest.mark.parametrize("mode", [O.REFERENCE, O.CONSISTENT])
def test_residual_of_constant_is_f(mode):
    n = 9
    s = O.Solver((n, n, n), 1, 1.0 / (n - 1), (0, 0, 0), "DDDDDD", mode, O.GAUSS_SEIDEL, O.INJECTION, 3)
    _fill_interior(s, "u", np.full((n, n, n), 0.5))
    s.residual(0)
    L = s.layout()
    np.testing.assert_array_equal(O.interior(L, s.array("r"))[1:-1, 1:-1, 1:-1],
                                  O.interior(L, s.array("f"))[1:-1, 1:-1, 1:-1])


@pytest.mark.parametrize("mode", [O.REFERENCE, O.CONSISTENT])
@pytest.mark.parametrize("abc", [(1, 1, 1), (2, 3, 1), (5, 1, 4)])
def test_jacobi_damps_sine_mode_by_eigenvalue(mode, abc):
    n = 17
    h = 1.0 / (n - 1)
    omega = 6.0 / 7.0
    s = O.Solver((n, n, n), 1, h, (0, 0, 0), "DDDDDD", mode, O.JACOBI, O.INJECTION, -1)
    X, Y, Z = _coords(s)
    a, b, c = abc
    m = np.sin(a * np.pi * X) * np.sin(b * np.pi * Y) * np.sin(c * np.pi * Z)
    _fill_interior(s, "u", m)
    s.relax(0, omega, 1)  # swap, then u = J(uprev)
    lam = 1 - omega * (1 - (np.cos(a * np.pi * h) + np.cos(b * np.pi * h) + np.cos(c * np.pi * h)) / 3)
    got = O.interior(s.layout(), s.array("u"))[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(got, lam * m[1:-1, 1:-1, 1:-1], atol=1e-14)


def _dense_solution(n, h, origin, bc, problem):
    """Independent statement of the discrete problem: 7-point Laplacian with
    ghost-node Neumann faces (ghost = mirror -+ 2h g) and Dirichlet values on
    the other faces, assembled as a dense matrix and solved directly."""
    L = O.lib()
    xs = [origin[a] + np.arange(n[a]) * h for a in range(3)]
    dirichlet = {q: bc[q] == "D" for q in range(6)}
    lo = [int(dirichlet[O.W]), int(dirichlet[O.S]), int(dirichlet[O.B])]
    hi = [n[0] - int(dirichlet[O.E]), n[1] - int(dirichlet[O.N]), n[2] - int(dirichlet[O.T])]
    unknown = [(i, j, k) for i in range(lo[0], hi[0]) for j in range(lo[1], hi[1]) for k in range(lo[2], hi[2])]
    pos = {p: m for m, p in enumerate(unknown)}
    A = np.zeros((len(unknown), len(unknown)))
    rhs = np.zeros(len(unknown))
    uf = lambda i, j, k: L.og_ufun(problem, xs[0][i], xs[1][j], xs[2][k])
    du = [L.og_dudx, L.og_dudy, L.og_dudz]
    for m, (i, j, k) in enumerate(unknown):
        rhs[m] = L.og_ffun(problem, xs[0][i], xs[1][j], xs[2][k]) * h * h
        A[m, m] -= 6.0
        p0 = [i, j, k]
        for ax in range(3):
            for d in (-1, 1):
                q = list(p0)
                q[ax] += d
                if 0 <= q[ax] < n[ax]:
                    t = tuple(q)
                    if t in pos:
                        A[m, pos[t]] += 1.0
                    else:  # a Dirichlet node
                        rhs[m] -= uf(*t)
                else:
                    # ghost: u(-1) = u(1) - 2h g (low face), u(n) = u(n-2) + 2h g
                    q[ax] -= 2 * d
                    g = du[ax](problem, xs[0][i], xs[1][j], xs[2][k])
                    t = tuple(q)
                    if t in pos:
                        A[m, pos[t]] += 1.0
                    else:
                        rhs[m] -= uf(*t)
                    rhs[m] -= d * 2 * h * g
    sol = np.linalg.solve(A, rhs)
    U = np.zeros(n)
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                U[i, j, k] = sol[pos[(i, j, k)]] if (i, j, k) in pos else uf(i, j, k)
    return U


@pytest.mark.parametrize("bc", ["DDDDDD", "NNNNDD", "DDNNND", "NDNDNN"])
def test_consistent_mode_converges_to_dense_discrete_solution(bc):
    n = (9, 9, 9)
    h = 1.0 / 8
    origin = (0.1, -0.2, 0.3)
    U = _dense_solution(n, h, origin, bc, 3)
    s = O.Solver(n, 3, h, origin, bc, O.CONSISTENT, O.GAUSS_SEIDEL, O.FULL_WEIGHTING, 3)
    s.to_device()
    # residual at the dense solution vanishes (pins the operator and the BCs)
    L = s.layout()
    u = s.array("u")
    keep = u.copy()
    O.interior(L, u)[...] = U
    s.residual(0)
    fmax = np.abs(s.array("f")).max()
    assert np.abs(O.interior(L, s.array("r"))).max() / fmax < 1e-12
    u[:] = keep
    it, fin, rr = s.solve(O.FCYCLE, 2, 1.0, maxiter=60, miniter=1, tol=1e-13)
    assert fin <= 1e-13
    np.testing.assert_allclose(O.interior(L, s.array("u")), U, rtol=0, atol=1e-11 * np.abs(U).max())


@pytest.mark.parametrize("relax", [O.GAUSS_SEIDEL, O.JACOBI])
def test_consistent_mode_second_order_convergence(relax):
    errs = []
    for n in (17, 33, 65):
        s = O.Solver((n, n, (n + 1) // 2), 4, 1.0 / (n - 1), (0, 0, 0), "NNNNDD", O.CONSISTENT, relax,
                     O.FULL_WEIGHTING, 2)
        s.to_device()
        it, fin, _ = s.solve(O.FCYCLE, 2, 1.0 if relax == O.GAUSS_SEIDEL else -1.0, maxiter=80, tol=1e-12)
        assert fin <= 1e-12
        X, Y, Z = _coords(s)
        errs.append(np.abs(O.interior(s.layout(), s.array("u")) - X ** 3 * Y ** 2 * Z).max())
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(3.5 < r < 4.5 for r in ratios), (errs, ratios)


def test_red_black_half_sweeps_zero_their_colour():
    """With w = 1 a red half-sweep solves every red equation exactly (its
    neighbours are black), so the residual vanishes on red points; black
    likewise.  Pins the colouring and the update formula."""
    n = 11
    s = O.Solver((n, n, n), 1, 0.1, (0, 0, 0), "DDDDDD", O.CONSISTENT, O.GAUSS_SEIDEL, O.INJECTION, 1)
    rng = np.random.default_rng(0)
    L = s.layout()
    O.interior(L, s.array("u"))[1:-1, 1:-1, 1:-1] = rng.standard_normal((n - 2,) * 3)
    I, J, K = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    inner = (I > 0) & (I < n - 1) & (J > 0) & (J < n - 1) & (K > 0) & (K < n - 1)
    fmax = np.abs(s.array("f")).max()
    for colour in (0, 1):
        O.lib().og_gs_half(O.CONSISTENT, ctypes.byref(L), O.ptr(s.array("u")), O.ptr(s.array("f")),
                           O.box_arr([1, n - 1, 1, n - 1, 1, n - 1]), 1.0, 0.1, colour)
        s.residual(0)
        r = O.interior(L, s.array("r"))
        sel = inner & (((I + J + K) % 2) == colour)
        assert np.abs(r[sel]).max() < 1e-10 * max(fmax, 1.0) / 0.01
        assert np.abs(r[inner & ~sel]).max() > 1e-6


def test_full_weighting_reproduces_constants_and_linears():
    Li = O.layout((17, 9, 13))
    Lo = O.layout((9, 5, 7))
    a = np.full(Li.size, 0.75)
    b = np.zeros(Lo.size)
    O.lib().og_restrict_fw(ctypes.byref(Li), O.ptr(a), ctypes.byref(Lo), O.ptr(b))
    np.testing.assert_array_equal(b, 0.75)
    I, J, K = np.meshgrid(np.arange(17), np.arange(9), np.arange(13), indexing="ij")
    a = np.ascontiguousarray((1.0 + 0.5 * I - 0.25 * J + 0.125 * K).astype(np.float64).ravel())
    O.lib().og_restrict_fw(ctypes.byref(Li), O.ptr(a), ctypes.byref(Lo), O.ptr(b))
    Ic, Jc, Kc = np.meshgrid(np.arange(9), np.arange(5), np.arange(7), indexing="ij")
    want = 1.0 + 0.5 * 2 * Ic - 0.25 * 2 * Jc + 0.125 * 2 * Kc
    b3 = b.reshape(9, 5, 7)
    np.testing.assert_array_equal(b3[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1])
    # clamped edge (fullweighting.h:71-79): at i = 0 the i-1 neighbour is i = 0
    # itself, weights 3/4 on x = 0 and 1/4 on x = h
    assert b3[0, 2, 3] == 0.75 * (1.0 - 0.25 * 4 + 0.125 * 6) + 0.25 * (1.5 - 0.25 * 4 + 0.125 * 6)


def test_injection_samples_even_points():
    Li = O.layout((9, 9, 9), (0, 1, 0, 1, 1, 0))
    Lo = O.layout((5, 5, 5), (0, 1, 0, 1, 1, 0))
    rng = np.random.default_rng(2)
    a = rng.standard_normal(Li.size)
    b = np.zeros(Lo.size)
    O.lib().og_restrict_inject(ctypes.byref(Li), O.ptr(a), ctypes.byref(Lo), O.ptr(b))
    np.testing.assert_array_equal(O.interior(Lo, b), O.interior(Li, a)[::2, ::2, ::2])


def test_trilinear_prolongation_exact_on_trilinear_functions():
    Li = O.layout((5, 3, 9), (1, 0, 1, 0, 0, 1))
    Lo = O.layout((9, 5, 17), (1, 0, 1, 0, 0, 1))
    f = lambda x, y, z: 0.5 + x - 2 * y + 0.25 * z + 0.5 * x * y - z * x + 0.125 * x * y * z
    I, J, K = np.meshgrid(np.arange(5.0), np.arange(3.0), np.arange(9.0), indexing="ij")
    a = np.zeros(Li.size)
    O.interior(Li, a)[...] = f(2 * I, 2 * J, 2 * K)
    b = np.zeros(Lo.size)
    O.lib().og_prolong_trilinear(ctypes.byref(Li), O.ptr(a), ctypes.byref(Lo), O.ptr(b))
    I, J, K = np.meshgrid(np.arange(9.0), np.arange(5.0), np.arange(17.0), indexing="ij")
    np.testing.assert_allclose(O.interior(Lo, b), f(I, J, K), rtol=0, atol=1e-12)


def test_neumann_ghost_is_exact_for_quadratics():
    """ghost = u(1) - 2h u'(0) equals u(-h) exactly when u is quadratic."""
    n = (7, 5, 5)
    h = 0.25
    halo = (0, 0, 1, 1, 0, 0)
    L = O.layout(n, halo)
    u = np.zeros(L.size)
    x = np.arange(n[0]) * h
    O.interior(L, u)[...] = (3 * x ** 2 - x + 2)[:, None, None]
    gW = np.full(n[1] * n[2], -1.0)  # u'(0)
    gE = np.full(n[1] * n[2], 6 * x[-1] - 1)
    O.lib().og_bc_neumann(ctypes.byref(L), O.ptr(u), O.W, O.ptr(gW), h)
    O.lib().og_bc_neumann(ctypes.byref(L), O.ptr(u), O.E, O.ptr(gE), h)
    u3 = u.reshape(O.padded_shape(L))
    q = lambda t: 3 * t ** 2 - t + 2
    np.testing.assert_allclose(u3[0], q(-h), rtol=1e-15)
    np.testing.assert_allclose(u3[-1], q(x[-1] + h), rtol=1e-15)


def test_norm_is_max_abs_over_shape_only():
    L = O.layout((5, 6, 7), (1, 1, 1, 1, 1, 1))
    rng = np.random.default_rng(3)
    a = rng.standard_normal(L.size)
    a3 = a.reshape(O.padded_shape(L))
    a3[0, 0, 0] = 100.0  # halo corner: must not count
    assert O.lib().og_norm_inf(ctypes.byref(L), O.ptr(a)) == np.abs(O.interior(L, a)).max()


def test_invalid_grid_is_rejected():
    with pytest.raises(ValueError):
        O.Solver((10, 9, 9), 3, 0.1)
