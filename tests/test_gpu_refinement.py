"""Mixed-precision paths on the GPU against the FP64 oracle.

* The hi + lo kernels (mgp_update_residual_hilo_f64_f32, mgp_residual_hilo_f64)
  and the FP64 coarsest-level sweeps of an FP32 level (mgp_rbgs_sweeps_f32_f64)
  equal the oracle bit for bit.
* Whole solves -- the all-Dirichlet FP32-cycle refinement solve of the
  1025x1025x513 mixed configuration (problem 3, at 129x129x65 and 257x257x129,
  with and without FP16 coarse levels), a Jacobi refinement solve, batched
  solves with 129-point axes and served requests with mixed faces -- agree with
  the oracle's FP64 solve of the same discrete problem to the bound their
  residuals allow: |u_a - u_b| <= ||Lap_h^-1|| (|r_a| + |r_b|) (tests/bounds.py).
* The served and benched right-hand side (mgp_fill_modes_f64) equals numpy's
  sum of sine products.
"""
import ctypes
import zlib

import numpy as np
import pytest

import bounds
import mgp
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def box_of(shape, bc):
    return [int(bc[O.W] == "D"), shape[0] - int(bc[O.E] == "D"), int(bc[O.S] == "D"), shape[1] - int(bc[O.N] == "D"),
            int(bc[O.B] == "D"), shape[2] - int(bc[O.T] == "D")]


def face_shape(n, q):
    m = list(n)
    m[{O.N: 1, O.S: 1, O.E: 0, O.W: 0, O.T: 2, O.B: 2}[q]] = 1
    return tuple(m)


def max_levels(d):
    """Levels down to a 3-point coarsest axis with exact 2:1 nesting."""
    L = 1
    while all(((n - 1) >> (L - 1)) % 2 == 0 and ((n - 1) >> (L - 1)) // 2 >= 2 for n in d):
        L += 1
    return L


def fourier_rhs(d, modes, h=None):
    """f = sum A sin(a pi x) sin(b pi y) sin(c pi z) at x = i h over the shape."""
    h = 1.0 / (d[0] - 1) if h is None else h
    x = np.arange(d[0])[:, None, None] * h
    y = np.arange(d[1])[None, :, None] * h
    z = np.arange(d[2])[None, None, :] * h
    f = np.zeros(d)
    for a, b, c, A in modes:
        f += A * np.sin(a * np.pi * x) * np.sin(b * np.pi * y) * np.sin(c * np.pi * z)
    return f


def random_modes(rng, count=4, kmax=8):
    return [(int(a), int(b), int(c), float(rng.standard_normal())) for a, b, c in rng.integers(1, kmax + 1, (count, 3))]


def oracle_solve(dims, levels, bc, relaxation, f=None, problem=-1, cycle=O.FCYCLE, nsmooth=2, omega=1.0,
                 coarse_sweeps=0, tol=1e-11, maxiter=80):
    h = 1.0 / (dims[0] - 1)
    o = O.Solver(dims, levels, h, (0, 0, 0), bc, O.CONSISTENT, relaxation, O.FULL_WEIGHTING, problem,
                 coarse_sweeps=coarse_sweeps)
    L = o.layout()
    if f is not None:
        O.interior(L, o.array("f"))[:] = f
    o.to_device()
    it, rel, _ = o.solve(cycle, nsmooth, omega, maxiter=maxiter, miniter=1, tol=tol)
    return o, it, rel, O.interior(L, o.array("u")).copy(), O.interior(L, o.array("f")).copy()


# ------------------------------------------------------------ kernels

HILO_SHAPES = [((40, 33, 130), "NNNNDD"), ((9, 17, 65), "DDDDDD"), ((33, 30, 64), "NDDNND"),
               ((17, 15, 129), "DNNDDN"), ((5, 5, 5), "DNDNDN"), ((65, 19, 70), "NNNNND")]


@pytest.mark.parametrize("rows", [16, 8])
@pytest.mark.parametrize("shape,bc", HILO_SHAPES, ids=[str(i) for i in range(len(HILO_SHAPES))])
def test_hilo_kernels_equal_oracle(shape, bc, rows):
    """hi + lo update and residual == the oracle bit for bit: u_new and lo_new
    over the shape (halos untouched), the FP32 residual and the FP64
    residual on the box, r outside the box untouched, and the norms."""
    say("case hilo kernels", shape, bc, rows)
    mgp.check(mgp.lib().mgp_tuning_set(b"update_residual.rows", rows))
    halo = [1 if c == "N" else 0 for c in bc]
    rng = np.random.default_rng(zlib.crc32(repr((shape, bc, "hilo")).encode()))
    ps = mgp.padded_shape(list(shape), halo)
    hi0 = rng.standard_normal(ps) * 2.0
    lo0 = (rng.standard_normal(ps) * 1e-16).astype(np.float32)
    e0 = (rng.standard_normal(ps) * 1e-3).astype(np.float32)
    f0, r0 = rng.standard_normal(ps) * 10.0, rng.standard_normal(ps)
    h = 1.0 / (shape[0] - 1)
    faces_np = {q: rng.standard_normal(int(np.prod(face_shape(shape, q)))) for q in range(6) if bc[q] == "N"}
    faces = {q: mgp.DeviceBuffer(a) for q, a in faces_np.items()}
    F = mgp.Faces()
    for q in range(6):
        F.type[q] = mgp.NEUMANN if bc[q] == "N" else mgp.DIRICHLET
        F.g[q] = faces[q].ptr if q in faces else None
    F.h = h
    box = box_of(shape, bc)
    # oracle
    L = O.layout(shape, halo)
    ohi, olo = hi0.ravel().copy(), lo0.ravel().copy()
    O.hilo_update(L, ohi, olo, e0.ravel().copy())
    orr = r0.ravel().copy()
    onrm = O.hilo_residual(L, ohi, olo, f0.ravel().copy(), orr, faces_np, box, h)
    # GPU: update + FP32 residual
    u = mgp.DeviceArray(shape, halo).upload(hi0)
    un = mgp.DeviceArray(shape, halo).upload(np.full(ps, 3.0))
    lo = mgp.DeviceArray(shape, halo, np.float32).upload(lo0)
    ln = mgp.DeviceArray(shape, halo, np.float32).upload(np.full(ps, 5.0, np.float32))
    e = mgp.DeviceArray(shape, halo, np.float32).upload(e0)
    f = mgp.DeviceArray(shape, halo).upload(f0)
    r = mgp.DeviceArray(shape, halo).upload(r0)
    out = mgp.DeviceArray(shape, halo, np.float32).upload(np.full(ps, 9.0, np.float32))
    nrm = mgp.DeviceBuffer(np.zeros(1))
    mgp.call("mgp_update_residual_hilo_f64_f32", u.ref(), un.ref(), lo.ref(), ln.ref(), e.ref(), f.ref(), r.ref(),
             out.ref(), ctypes.byref(F), mgp.box(box), nrm.ptr, None)
    inner = O.interior(L, np.arange(L.size)).ravel()
    gun, gln = un.download().ravel(), ln.download().ravel()
    np.testing.assert_array_equal(gun[inner], ohi[inner])
    np.testing.assert_array_equal(gln[inner], olo[inner])
    halo_pts = np.setdiff1d(np.arange(L.size), inner)
    assert np.all(gun[halo_pts] == 3.0) and np.all(gln[halo_pts] == 5.0)
    gout = out.download().ravel()
    want = np.full(L.size, 9.0, np.float32)
    bmask = np.zeros(ps, bool)
    bi = O.interior(L, bmask.ravel())
    bi[box[0]:box[1], box[2]:box[3], box[4]:box[5]] = True
    want[bmask.ravel()] = orr[bmask.ravel()].astype(np.float32)
    np.testing.assert_array_equal(gout, want)
    assert nrm.download()[0] == onrm
    np.testing.assert_array_equal(u.download(), hi0)  # inputs not written
    # GPU: FP64 residual of the updated iterate
    r2 = mgp.DeviceArray(shape, halo).upload(r0)
    nrm2 = mgp.DeviceBuffer(np.zeros(1))
    mgp.call("mgp_residual_hilo_f64", un.ref(), ln.ref(), f.ref(), r2.ref(), ctypes.byref(F), mgp.box(box), nrm2.ptr,
             None)
    np.testing.assert_array_equal(r2.download().ravel(), orr)
    assert nrm2.download()[0] == onrm
    # rejected: in-place update, mismatched layouts
    lib = mgp.lib()
    assert lib.mgp_update_residual_hilo_f64_f32(u.ref(), u.ref(), lo.ref(), ln.ref(), e.ref(), f.ref(), r.ref(),
                                                out.ref(), ctypes.byref(F), mgp.box(box), None, None) == mgp.ERR_ARG
    other = mgp.DeviceArray((shape[0], shape[1], shape[2] + 1), halo, np.float32)
    assert lib.mgp_update_residual_hilo_f64_f32(u.ref(), un.ref(), lo.ref(), other.ref(), e.ref(), f.ref(), r.ref(),
                                                out.ref(), ctypes.byref(F), mgp.box(box), None, None) == mgp.ERR_ARG
    mgp.check(mgp.lib().mgp_tuning_set(b"update_residual.rows", 16))


COARSE_SHAPES = [((17, 17, 9), "DDDDDD"), ((9, 9, 9), "NNNNDD"), ((17, 9, 5), "DNDNDN"), ((5, 5, 5), "NNNNND")]


@pytest.mark.parametrize("shape,bc", COARSE_SHAPES, ids=[str(i) for i in range(len(COARSE_SHAPES))])
def test_fp64_coarsest_sweeps_equal_oracle(shape, bc):
    """mgp_rbgs_sweeps_f32_f64 (FP32 arrays, FP64 arithmetic) == the oracle's
    FP64 sweeps (ghost refresh + red + ghost refresh + black, consistent mode)
    on the widened FP32 values, rounded once to FP32, bit for bit."""
    say("case fp64 coarsest", shape, bc)
    halo = [1 if c == "N" else 0 for c in bc]
    rng = np.random.default_rng(zlib.crc32(repr((shape, bc, "coarse64")).encode()))
    ps = mgp.padded_shape(list(shape), halo)
    u0 = rng.standard_normal(ps).astype(np.float32)
    f0 = (rng.standard_normal(ps) * 50).astype(np.float32)
    h = 1.0 / (shape[0] - 1)
    faces_np = {q: rng.standard_normal(int(np.prod(face_shape(shape, q)))).astype(np.float32)
                for q in range(6) if bc[q] == "N"}
    faces = {q: mgp.DeviceBuffer(a) for q, a in faces_np.items()}
    F = mgp.Faces()
    for q in range(6):
        F.type[q] = mgp.NEUMANN if bc[q] == "N" else mgp.DIRICHLET
        F.g[q] = faces[q].ptr if q in faces else None
    F.h = h
    box = box_of(shape, bc)
    nsweeps, omega = 7, 1.0
    L = O.layout(shape, halo)
    ou = u0.astype(np.float64).ravel().copy()
    of = f0.astype(np.float64).ravel().copy()
    ofaces = {q: a.astype(np.float64) for q, a in faces_np.items()}
    lib = O.lib()
    for _ in range(nsweeps):
        for colour in (0, 1):
            for q in (O.E, O.W, O.N, O.S, O.T, O.B):
                if q in ofaces:
                    lib.og_bc_neumann(ctypes.byref(L), O.ptr(ou), q, O.ptr(ofaces[q]), h)
            lib.og_gs_half(O.CONSISTENT, ctypes.byref(L), O.ptr(ou), O.ptr(of), ctypes.byref(O.box_arr(box)), omega,
                           h, colour)
    want = O.interior(L, ou).astype(np.float32)
    a = mgp.DeviceArray(shape, halo, np.float32).upload(u0)
    b = mgp.DeviceArray(shape, halo, np.float32).upload(u0)
    f = mgp.DeviceArray(shape, halo, np.float32).upload(f0)
    mgp.call("mgp_rbgs_sweeps_f32_f64", mgp.CONSISTENT, a.ref(), b.ref(), f.ref(), ctypes.byref(F), mgp.box(box),
             omega, nsweeps, None)
    mgp.call("mgp_rbgs_sweeps_f32_f64", mgp.CONSISTENT, a.ref(), a.ref(), f.ref(), ctypes.byref(F), mgp.box(box),
             omega, nsweeps, None)
    np.testing.assert_array_equal(O.interior(L, b.download().ravel()), want)
    np.testing.assert_array_equal(O.interior(L, a.download().ravel()), want)
    # differs from the FP32-arithmetic sweeps (the arithmetic is FP64)
    c = mgp.DeviceArray(shape, halo, np.float32).upload(u0)
    mgp.call("mgp_rbgs_sweeps_f32", mgp.CONSISTENT, c.ref(), c.ref(), f.ref(), ctypes.byref(F), mgp.box(box),
             omega, nsweeps, None)
    assert not np.array_equal(O.interior(L, c.download().ravel()), want)


def test_fill_modes_matches_numpy():
    """The served/benched right-hand side: sum A sin(a pi x) sin(b pi y)
    sin(c pi z) at x = i h over the shape (halo untouched), against numpy's
    sines to a few ulp of sum |A|."""
    rng = np.random.default_rng(zlib.crc32(b"modes"))
    shape, halo = (33, 17, 9), (1, 0, 1, 0, 0, 1)
    modes = random_modes(rng, 7)
    h = 1.0 / 32
    f = mgp.DeviceArray(shape, halo).upload(np.full(mgp.padded_shape(list(shape), halo), 4.0))
    abc = (ctypes.c_int32 * (3 * len(modes)))(*[v for m in modes for v in m[:3]])
    amp = (ctypes.c_double * len(modes))(*[m[3] for m in modes])
    mgp.call("mgp_fill_modes_f64", f.ref(), len(modes), abc, amp, h, None)
    L = O.layout(shape, halo)
    got = f.download().ravel()
    want = fourier_rhs(shape, modes, h)
    tol = 4 * np.spacing(sum(abs(m[3]) for m in modes))
    np.testing.assert_allclose(O.interior(L, got), want, rtol=0, atol=tol)
    assert np.abs(O.interior(L, got)).max() > 0.1
    inner = O.interior(L, np.arange(L.size)).ravel()
    assert np.all(np.delete(got, inner) == 4.0)
    # sin(0) = 0 exactly on the x = 0, y = 0 and z = 0 planes
    g3 = O.interior(L, got)
    assert np.all(g3[0] == 0) and np.all(g3[:, 0] == 0) and np.all(g3[:, :, 0] == 0)
    assert mgp.lib().mgp_fill_modes_f64(f.ref(), 17, abc, amp, h, None) == mgp.ERR_ARG


# ------------------------------------------------------------ whole solves

def mixed_solver(dims, levels, bc, relaxation, problem, half_from=0, flags=mgp.QUIET, tol=1e-10, cs=16, maxiter=60):
    h = 1.0 / (dims[0] - 1)
    s = mgp.Solver(dims, levels, h, (0, 0, 0), bc, mgp.CONSISTENT, relaxation, mgp.FULL_WEIGHTING, problem,
                   dtype=mgp.MIXED, flags=flags, maxiter=maxiter, miniter=1, tolerance=tol, coarse_sweeps=cs,
                   half_from=half_from)
    s.to_device()
    return s


def solution(s, L):
    s.to_host()
    return O.interior(L, np.array(s.array("u"), copy=True).ravel())


P3_CASES = [((129, 129, 65), 5, 0), ((129, 129, 65), 5, 3), ((257, 257, 129), 6, 0), ((257, 257, 129), 6, 3)]


@pytest.mark.parametrize("case", P3_CASES, ids=[str(i) for i in range(len(P3_CASES))])
def test_mixed_dirichlet_problem3_matches_oracle(case):
    """The mixed configuration's problem (u = cos(x z^2) sin(y^3), all faces
    Dirichlet, FP32 F-cycles inside FP64 refinement, hi + lo iterate, FP64
    coarsest level, optionally FP16 levels from 129x129x65's level 3 down)
    reaches ||r||/||f|| <= 1e-10 and equals the oracle's FP64 F-cycle solve
    within the residual bound; both are the discretisation error away from
    the analytic u."""
    dims, levels, half_from = case
    say("case mixed P3", case)
    bc = "DDDDDD"
    s = mixed_solver(dims, levels, bc, mgp.GAUSS_SEIDEL, 3, half_from=half_from)
    it, rel, _ = s.solve("Fcycle", 2, 1.0)
    o, oit, orel, ou, of = oracle_solve(dims, levels, bc, O.GAUSS_SEIDEL, problem=3, coarse_sweeps=16, tol=1e-10)
    L = o.layout()
    u = solution(s, L)
    say("  mixed %d iterations rel %.3e, oracle %d cycles rel %.3e" % (it, rel, oit, orel))
    assert rel <= 1e-10 and orel <= 1e-10
    assert it <= oit + 2
    h = 1.0 / (dims[0] - 1)
    fnorm = np.abs(of).max()
    gap = np.abs(u - ou).max()
    bound = bounds.gap_bound(dims, h, bc, rel, orel, fnorm, umax=np.abs(ou).max())
    say("  |u - u_oracle| %.3e, bound %.3e" % (gap, bound))
    assert gap <= bound
    X, Y, Z = np.meshgrid(*[np.arange(n) * h for n in dims], indexing="ij")
    exact = np.cos(X * Z * Z) * np.sin(Y ** 3)
    err, oerr = np.abs(u - exact).max(), np.abs(ou - exact).max()
    say("  error vs analytic: mixed %.6e oracle %.6e" % (err, oerr))
    assert abs(err - oerr) <= bound
    s.close()


def test_mixed_hilo_passes_the_plain_fp64_floor():
    """At 257x257x129 the plain FP64 iterate stalls where 12/h^2 ulp(u)
    dominates the residual; the hi + lo iterate keeps converging to 1e-12,
    and its lo word is at most half an ulp of hi."""
    dims, levels = (257, 257, 129), 6
    say("case hilo floor")
    s = mixed_solver(dims, levels, "DDDDDD", mgp.GAUSS_SEIDEL, 3, tol=1e-12, maxiter=40)
    it, rel, _ = s.solve("Fcycle", 2, 1.0)
    p = mixed_solver(dims, levels, "DDDDDD", mgp.GAUSS_SEIDEL, 3, tol=1e-12, maxiter=40,
                     flags=mgp.QUIET | mgp.PLAIN_ITERATE)
    pit, prel, _ = p.solve("Fcycle", 2, 1.0)
    say("  hi+lo: %d iterations rel %.3e; plain: %d iterations rel %.3e" % (it, rel, pit, prel))
    assert rel <= 1e-12
    assert prel > rel  # the plain iterate ends at its floor (stagnation stop)
    s.to_host()
    hi = np.array(s.array("u"), copy=True).ravel()
    lo = s.lo_word().astype(np.float64)
    assert np.all(np.abs(lo) <= np.spacing(np.abs(hi)) / 2 * (1 + 2.0 ** -20))
    assert np.abs(lo).max() > 0
    s.close()
    p.close()


def test_mixed_jacobi_matches_oracle():
    """FP32 Jacobi V-cycles (omega 6/7, 3+3; the batched configuration's
    smoother, include/jacobi.h:59-81) inside FP64 refinement against the
    oracle's FP64 Jacobi V-cycle solve."""
    dims, levels, bc = (65, 65, 65), 5, "DDDDDD"
    say("case mixed jacobi")
    s = mixed_solver(dims, levels, bc, mgp.JACOBI, 3, cs=0)
    it, rel, _ = s.solve("Vcycle", 3, 6.0 / 7.0)
    o, oit, orel, ou, of = oracle_solve(dims, levels, bc, O.JACOBI, problem=3, cycle=O.VCYCLE, nsmooth=3,
                                        omega=6.0 / 7.0, tol=1e-10)
    u = solution(s, o.layout())
    say("  mixed %d rel %.3e, oracle %d rel %.3e" % (it, rel, oit, orel))
    assert rel <= 1e-10 and orel <= 1e-10
    bound = bounds.gap_bound(dims, 1.0 / 64, bc, rel, orel, np.abs(of).max(), umax=np.abs(ou).max())
    assert np.abs(u - ou).max() <= bound
    s.close()


BATCH_SHAPES = [(129, 65, 33), (33, 129, 65), (65, 33, 129), (129, 129, 33), (33, 33, 33), (65, 65, 65),
                (33, 65, 129), (129, 33, 65)]


def test_batch_with_129_axes_matches_oracle():
    """A batch of the batched configuration's kind (zero Dirichlet, four
    Fourier modes, Jacobi 6/7 with 3+3 sweeps, FP32 V-cycles in FP64
    refinement) with 129-point axes: each problem's solution against the
    oracle's FP64 Jacobi V-cycle solve of the same problem."""
    rng = np.random.default_rng(zlib.crc32(b"batch-oracle"))
    fs = [fourier_rhs(d, random_modes(rng)) for d in BATCH_SHAPES]
    b = mgp.Batch(BATCH_SHAPES, relaxation=mgp.JACOBI, nsmooth=3, omega=6.0 / 7.0, maxiter=40, tolerance=1e-10)
    for p, f in enumerate(fs):
        b.set_rhs(p, f)
    b.solve()
    for p, (d, f) in enumerate(zip(BATCH_SHAPES, fs)):
        it, rel = b.stats(p)
        u = b.solution(p)
        say("case batch-oracle", d, it, "%.3e" % rel)
        o, oit, orel, ou, of = oracle_solve(d, max_levels(d), "DDDDDD", O.JACOBI, f=f, cycle=O.VCYCLE, nsmooth=3,
                                            omega=6.0 / 7.0, tol=1e-11)
        assert rel <= 1e-10 and orel <= 1e-11
        bound = bounds.gap_bound(d, 1.0 / (d[0] - 1), "DDDDDD", rel, orel, np.abs(f).max(), umax=np.abs(ou).max())
        gap = np.abs(u - ou).max()
        say("  |u - u_oracle| %.3e bound %.3e" % (gap, bound))
        assert gap <= bound
        o.close()
    b.close()


def served_requests(seed, count, kmin=5, kmax=7):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        dims = tuple(int(2 ** k + 1) for k in rng.integers(kmin, kmax + 1, size=3))
        bc = "".join("N" if v else "D" for v in rng.integers(0, 2, size=6))
        if "D" not in bc:
            continue
        out.append((dims, bc, random_modes(rng)))
    return out


def test_served_requests_match_oracle():
    """Requests of the serving configuration's kind (2^k+1 per axis, k up to
    7, face types by seed, Fourier right-hand side, tolerance 1e-8), served
    four at a time: each result against the oracle's FP64 F-cycle solve of
    the same problem (its right-hand side from numpy's sines)."""
    reqs = served_requests(zlib.crc32(b"serve-oracle"), 9)
    s = mgp.Server(max_active=4, tolerance=1e-8)
    ids = [s.submit(dims, bc, modes) for dims, bc, modes in reqs]
    for rid, (dims, bc, modes) in zip(ids, reqs):
        res, u = s.wait(rid)
        say("case served-oracle", dims, bc, res.iterations, "%.3e" % res.rel_res)
        assert res.status == 0 and res.rel_res <= 1e-8
        f = fourier_rhs(dims, modes)
        o, oit, orel, ou, of = oracle_solve(dims, max_levels(dims), bc, O.GAUSS_SEIDEL, f=f, coarse_sweeps=32,
                                            tol=1e-11)
        assert orel <= 1e-11
        df = 4 * np.spacing(sum(abs(m[3]) for m in modes))  # device sines vs numpy's
        bound = bounds.gap_bound(dims, 1.0 / (dims[0] - 1), bc, res.rel_res, orel, np.abs(f).max(), df=df,
                                 umax=np.abs(ou).max())
        gap = np.abs(u - ou).max()
        say("  |u - u_oracle| %.3e bound %.3e" % (gap, bound))
        assert gap <= bound
        o.close()
    s.close()


def test_server_zero_omega_means_default():
    """omega = 0 in the server configuration selects the smoother's default
    (8/9 for Gauss-Seidel) instead of a smoother that does nothing."""
    reqs = served_requests(zlib.crc32(b"omega0"), 1, 4, 5)
    dims, bc, modes = reqs[0]
    s = mgp.Server(max_active=1, omega=0.0, tolerance=1e-8)
    res, u = s.wait(s.submit(dims, bc, modes))
    s.close()
    d = mgp.Server(max_active=1, omega=8.0 / 9.0, tolerance=1e-8)
    res2, u2 = d.wait(d.submit(dims, bc, modes))
    d.close()
    assert res.status == 0 and res.rel_res <= 1e-8 and res.iterations < 30
    assert (res.iterations, res.rel_res) == (res2.iterations, res2.rel_res)
    np.testing.assert_array_equal(u, u2)
