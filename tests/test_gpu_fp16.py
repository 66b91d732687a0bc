"""FP16 storage of the coarse levels of the FP32 cycle, on the GPU.

Kernels (mgp_*_f16 through the C ABI) equal the oracle's FP16-storage
operators bit for bit on random inputs; the LDS multi-sweep equals the
per-colour half-sweeps; a refinement solve with levels >= half_from in FP16
reaches the same FP64 tolerance as the all-FP32 cycle in at most one more
iteration, and one iteration's result stays within the FP16 rounding budget
of the all-FP32 one.
"""
import sys
import zlib

import numpy as np
import pytest

import bounds
import mgp
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


NOHALO = (0, 0, 0, 0, 0, 0)
SHAPES = [(5, 5, 5), (9, 9, 9), (17, 9, 33), (33, 33, 17), (65, 33, 33), (129, 65, 65)]


def rng_for(*key):
    return np.random.default_rng(zlib.crc32(repr(key).encode()))


def dev16(a):
    return mgp.DeviceArray(a.shape, NOHALO, dtype=np.float16).upload(a)


def dev32(a):
    return mgp.DeviceArray(a.shape, NOHALO, dtype=np.float32).upload(a)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint16 if a.dtype == np.float16 else np.uint32)


def interior_box(shape):
    return [1, shape[0] - 1, 1, shape[1] - 1, 1, shape[2] - 1]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("colour", [0, 1])
def test_half_sweep_f16_equals_oracle(shape, colour):
    say("case f16 half-sweep", shape, colour)
    rng = rng_for("gs", shape, colour)
    u = rng.standard_normal(shape).astype(np.float16)
    f = (8 * rng.standard_normal(shape)).astype(np.float16)
    box = interior_box(shape)
    h, omega = 1.0 / (shape[0] - 1), 8.0 / 9.0
    du, df = dev16(u), dev16(f)
    mgp.call("mgp_rbgs_half_f16", du.ref(), df.ref(), mgp.box(box), omega, h, colour, None)
    O.h_gs_half(u, f, box, omega, h, colour)
    assert np.array_equal(bits(du.download()), bits(u))


@pytest.mark.parametrize("shape", [s for s in SHAPES if np.prod(s) <= 8192] + [(17, 17, 25)])
@pytest.mark.parametrize("nsweeps", [1, 3])
def test_lds_sweeps_f16_equal_half_sweeps_and_oracle(shape, nsweeps):
    say("case f16 lds sweeps", shape, nsweeps)
    rng = rng_for("lds", shape, nsweeps)
    u = rng.standard_normal(shape).astype(np.float16)
    f = (8 * rng.standard_normal(shape)).astype(np.float16)
    box = interior_box(shape)
    h, omega = 1.0 / (shape[0] - 1), 8.0 / 9.0
    a, b, df = dev16(u), dev16(u), dev16(f)
    mgp.call("mgp_rbgs_sweeps_f16", a.ref(), df.ref(), mgp.box(box), omega, h, nsweeps, None)
    for _ in range(nsweeps):
        for c in (0, 1):
            mgp.call("mgp_rbgs_half_f16", b.ref(), df.ref(), mgp.box(box), omega, h, c, None)
            O.h_gs_half(u, f, box, omega, h, c)
    got = a.download()
    assert np.array_equal(bits(got), bits(b.download()))
    assert np.array_equal(bits(got), bits(u))


@pytest.mark.parametrize("shape", SHAPES)
def test_residual_f16_equals_oracle(shape):
    say("case f16 residual", shape)
    rng = rng_for("res", shape)
    u = rng.standard_normal(shape).astype(np.float16)
    f = (8 * rng.standard_normal(shape)).astype(np.float16)
    r = np.zeros(shape, dtype=np.float16)
    r[0] = 0.25  # a stored value outside the box counts in the norm
    box = interior_box(shape)
    h = 1.0 / (shape[0] - 1)
    du, df, dr = dev16(u), dev16(f), dev16(r)
    nrm = mgp.DeviceBuffer(np.zeros(1))
    mgp.call("mgp_residual_f16", du.ref(), df.ref(), dr.ref(), mgp.box(box), h, nrm.ptr, None)
    m = O.h_residual(u, f, r, box, h)
    assert np.array_equal(bits(dr.download()), bits(r))
    assert nrm.download()[0] == m


@pytest.mark.parametrize("fine", [(9, 9, 9), (33, 17, 65), (129, 65, 65), (65, 129, 33)])
@pytest.mark.parametrize("src", ["f16", "f32"])
@pytest.mark.parametrize("norm", [0.0, 0.37, 5e-9, 3.0e4])
def test_restrict_fw_scaled_equals_oracle(fine, src, norm):
    say("case f16 fw", fine, src, norm)
    rng = rng_for("fw", fine, src)
    a = (norm or 1.0) * rng.uniform(-1, 1, fine)
    a = a.astype(np.float16 if src == "f16" else np.float32)
    coarse = tuple((n - 1) // 2 + 1 for n in fine)
    c = np.zeros(coarse, dtype=np.float16)
    da = dev16(a) if src == "f16" else dev32(a)
    dc = dev16(c)
    nrm = mgp.DeviceBuffer(np.array([norm]))
    mgp.call("mgp_restrict_fw_scaled_" + ("f16" if src == "f16" else "f32_f16"), da.ref(), dc.ref(), nrm.ptr, None)
    O.h_restrict_fw(a, c, norm)
    got = dc.download()
    assert np.array_equal(bits(got), bits(c))
    amax = np.abs(a.astype(np.float64)).max()
    if 0 < amax <= norm:  # norm bounds the (non-zero) input: output scaled towards 1
        assert 0.05 < np.abs(got.astype(np.float64)).max() < 1.0


@pytest.mark.parametrize("coarse", [(5, 5, 5), (17, 9, 33), (65, 33, 33), (33, 65, 17)])
@pytest.mark.parametrize("dst", ["f16", "f32"])
@pytest.mark.parametrize("norm", [0.0, 0.37, 5e-9])
def test_prolong_add_scaled_equals_oracle(coarse, dst, norm):
    say("case f16 prolong", coarse, dst, norm)
    rng = rng_for("pr", coarse, dst)
    c = rng.uniform(-1, 1, coarse).astype(np.float16)
    fine = tuple(2 * (n - 1) + 1 for n in coarse)
    u = ((norm or 1.0) * rng.standard_normal(fine)).astype(np.float16 if dst == "f16" else np.float32)
    dc = dev16(c)
    du = dev16(u) if dst == "f16" else dev32(u)
    nrm = mgp.DeviceBuffer(np.array([norm]))
    mgp.call("mgp_prolong_add_scaled_f16" + ("" if dst == "f16" else "_f32"), dc.ref(), du.ref(), nrm.ptr, None)
    O.h_prolong_add(c, u, norm)
    assert np.array_equal(bits(du.download()), bits(u))


def test_f16_kernels_reject_bad_arguments():
    L = mgp.lib()
    u = dev16(np.zeros((9, 9, 9), dtype=np.float16))
    f = dev16(np.zeros((9, 9, 9), dtype=np.float16))
    big = dev16(np.zeros((33, 33, 17), dtype=np.float16))
    bigf = dev16(np.zeros((33, 33, 17), dtype=np.float16))
    c_bad = dev16(np.zeros((4, 5, 5), dtype=np.float16))
    nrm = mgp.DeviceBuffer(np.zeros(1))
    touching = [0, 9, 1, 8, 1, 8]  # touches the west face: Dirichlet faces only
    assert L.mgp_rbgs_half_f16(u.ref(), f.ref(), mgp.box(touching), 1.0, 0.1, 0, None) == mgp.ERR_ARG
    assert L.mgp_rbgs_half_f16(u.ref(), f.ref(), mgp.box(interior_box((9, 9, 9))), 1.0, 0.1, 2, None) == mgp.ERR_ARG
    assert L.mgp_rbgs_sweeps_f16(big.ref(), bigf.ref(), mgp.box(interior_box((33, 33, 17))), 1.0, 0.1, 1,
                                 None) == mgp.ERR_ARG  # 18513 points: more than one workgroup's LDS
    assert L.mgp_residual_f16(u.ref(), big.ref(), u.ref(), mgp.box(interior_box((9, 9, 9))), 0.1, None,
                              None) == mgp.ERR_ARG
    assert L.mgp_restrict_fw_scaled_f16(u.ref(), c_bad.ref(), nrm.ptr, None) == mgp.ERR_ARG
    assert L.mgp_restrict_fw_scaled_f16(u.ref(), c_bad.ref(), None, None) == mgp.ERR_ARG
    assert L.mgp_prolong_add_scaled_f16(c_bad.ref(), u.ref(), nrm.ptr, None) == mgp.ERR_ARG


# --------------------------------------------------------- solver level

def solver(dims, levels, half_from, flags=mgp.QUIET, maxiter=40, tol=1e-10, problem=3):
    h = 1.0 / (dims[0] - 1)
    s = mgp.Solver(dims, levels, h, (0, 0, 0), "DDDDDD", mgp.CONSISTENT, mgp.GAUSS_SEIDEL, mgp.FULL_WEIGHTING,
                   problem, dtype=mgp.MIXED, flags=flags, maxiter=maxiter, miniter=1, tolerance=tol,
                   coarse_sweeps=16, half_from=half_from)
    s.to_device()
    return s


def solution(s):
    s.to_host()
    host, a = s._arr(0, "u")
    return np.array(s.array("u"), copy=True)


FP16_CASES = [((65, 65, 33), 4, 2), ((129, 129, 65), 5, 3), ((129, 129, 65), 5, 2), ((257, 257, 129), 6, 3)]


@pytest.mark.parametrize("case", FP16_CASES, ids=[str(i) for i in range(len(FP16_CASES))])
@pytest.mark.parametrize("cyc", ["Fcycle", "Vcycle"])
def test_fp16_levels_refinement_reaches_tolerance(case, cyc):
    """Levels >= half_from in FP16: ||r||/||f|| <= 1e-10 of the FP64 operator
    within one iteration more than the all-FP32 cycle, the same discrete
    solution to 1e-8, and graph replay equal to launch-by-launch, bit for
    bit."""
    dims, levels, half_from = case
    say("case fp16 solve", case, cyc)
    ref = solver(dims, levels, 0)
    it32, rel32, _ = ref.solve(cyc, 2, 1.0)
    u32 = solution(ref)
    ref.close()
    runs = {}
    for name, flags in (("graphs", mgp.QUIET), ("nographs", mgp.QUIET | mgp.NO_GRAPHS)):
        s = solver(dims, levels, half_from, flags)
        it, rel, _ = s.solve(cyc, 2, 1.0)
        runs[name] = (it, rel, solution(s))
        s.close()
    it, rel, u16 = runs["graphs"]
    say("  iterations fp32 %d fp16-coarse %d, rel %.3e / %.3e" % (it32, it, rel32, rel))
    assert rel <= 1e-10 and rel32 <= 1e-10
    assert it <= it32 + 1, (it, it32)
    # both solve one discrete problem: the gap their residuals allow
    s0 = solver(dims, levels, 0, maxiter=0)
    fnorm = np.abs(np.array(s0.array("f"), copy=True)).max()
    s0.close()
    bound = bounds.gap_bound(dims, 1.0 / (dims[0] - 1), "DDDDDD", rel, rel32, fnorm, umax=np.abs(u32).max())
    say("  |u16 - u32| %.3e, bound %.3e" % (np.max(np.abs(u16 - u32)), bound))
    assert np.max(np.abs(u16 - u32)) <= bound
    assert runs["nographs"][:2] == runs["graphs"][:2]
    assert np.array_equal(runs["nographs"][2], u16)


@pytest.mark.parametrize("half_from", [2, 3, 4])
def test_fp16_levels_first_iteration_within_rounding_budget(half_from):
    """One refinement iteration (u = 0 + e, e one F-cycle) with FP16 levels
    differs from the all-FP32 iteration by far less than the error that
    iteration leaves: the FP16 levels perturb the coarse-grid correction by
    their rounding (2^-11 relative per stored value), they do not change it."""
    dims, levels = (129, 129, 65), 5
    say("case fp16 first iteration", half_from)
    a = solver(dims, levels, 0, maxiter=1)
    a.solve("Fcycle", 2, 1.0)
    ua = solution(a)
    a.close()
    b = solver(dims, levels, half_from, maxiter=1)
    b.solve("Fcycle", 2, 1.0)
    ub = solution(b)
    b.close()
    c = solver(dims, levels, 0, maxiter=40)
    c.solve("Fcycle", 2, 1.0)
    uc = solution(c)  # the discrete solution (to 1e-10)
    c.close()
    diff = np.max(np.abs(ub - ua))
    err = np.max(np.abs(ua - uc))
    say("  half_from %d: |u1_fp16 - u1_fp32| %.3e, error after one iteration %.3e, ratio %.2e"
        % (half_from, diff, err, diff / err))
    assert 0.0 < diff <= 0.05 * err


def test_fp16_levels_reject_unsupported_settings():
    h = 1.0 / 64
    common = dict(mode=mgp.CONSISTENT, relaxation=mgp.GAUSS_SEIDEL, restriction=mgp.FULL_WEIGHTING, problem=3)
    with pytest.raises(mgp.MgpError):
        mgp.Solver((65, 65, 33), 4, h, (0, 0, 0), "NNNNDD", dtype=mgp.MIXED, half_from=2, **common)
    with pytest.raises(mgp.MgpError):
        mgp.Solver((65, 65, 33), 4, h, (0, 0, 0), "DDDDDD", dtype=mgp.MIXED, half_from=4, **common)
    with pytest.raises(mgp.MgpError):
        mgp.Solver((65, 65, 33), 4, h, (0, 0, 0), "DDDDDD", dtype=mgp.MIXED, half_from=1, **common)
    with pytest.raises(mgp.MgpError):
        mgp.Solver((65, 65, 33), 4, h, (0, 0, 0), "DDDDDD", dtype=mgp.F32, half_from=2, **common)
    with pytest.raises(mgp.MgpError):
        mgp.Solver((65, 65, 33), 4, h, (0, 0, 0), "DDDDDD", mode=mgp.CONSISTENT, relaxation=mgp.JACOBI,
                   restriction=mgp.FULL_WEIGHTING, problem=3, dtype=mgp.MIXED, half_from=2)
