"""mgp_cycle_noalloc_f64 (cycles without device allocation, the coarse levels
in the library's static scratch array) against the existing entry point it
varies, mgp_solver_cycle on a consistent-mode FP64 Gauss-Seidel /
full-weighting solver: the same u bit for bit and the same residual norm, on
two calls in a row, the second on a smaller problem (the scratch array keeps
the first call's contents)."""
import ctypes
import sys

import numpy as np
import pytest

import mgp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_copy(desc):
    """A DeviceArray holding the bytes of a solver's device array."""
    shape = [desc.shape[0], desc.shape[1], desc.shape[2]]
    halo = [desc.halo[q] for q in range(6)]
    a = mgp.DeviceArray(shape, halo)
    host = np.empty(a.pshape)
    mgp.check(mgp.lib().mgp_stream_sync(None))
    mgp.check(mgp.lib().mgp_memcpy_d2h(host.ctypes.data, desc.data, a.nbytes, None))
    mgp.check(mgp.lib().mgp_stream_sync(None))
    return a.upload(host)


def interior(a, desc):
    hN, hS, hE, hW, hT, hB = [desc.halo[q] for q in range(6)]
    n = desc.shape
    return a[hW:hW + n[0], hS:hS + n[1], hB:hB + n[2]]


CASES = [((65, 65, 33), 4, "NNNNDD", 2, "Fcycle", 2, 2, 1.0, 8),
         ((33, 33, 17), 3, "DDDDDD", 3, "Vcycle", 3, 3, -1.0, 0),
         ((17, 33, 9), 3, "NDNDND", 2, "Fcycle", 1, 2, 1.0, 6)]


def run_case(case):
    dims, levels, bc, problem, cyc, ncycles, nsmooth, omega, cs = case
    say("case noalloc", case)
    h = 1.0 / (dims[0] - 1)
    L = mgp.lib()
    s = mgp.Solver(dims, levels, h, (0.0, 0.0, 0.0), bc, mgp.CONSISTENT, mgp.GAUSS_SEIDEL, mgp.FULL_WEIGHTING,
                   problem, dtype=mgp.F64, maxiter=100, tolerance=0.0, coarse_sweeps=cs)
    s.to_device()
    ud, fd = s.device_array("u"), s.device_array("f")
    u = device_copy(ud)
    f = device_copy(fd)
    F = mgp.Faces()
    for q in range(6):
        host, a, t = mgp.vp(), mgp.Array(), ctypes.c_int()
        mgp.check(L.mgp_solver_face(s.h, 0, q, ctypes.byref(host), ctypes.byref(a), ctypes.byref(t)))
        F.type[q] = t.value
        F.g[q] = a.data if t.value == mgp.NEUMANN else None
    F.h = h
    nrm = mgp.DeviceBuffer(np.array([123.0]))
    rc = L.mgp_cycle_noalloc_f64(cyc.encode(), u.ref(), f.ref(), ctypes.byref(F), levels, ncycles, nsmooth, omega, cs,
                                 ctypes.c_void_p(nrm.ptr), None)
    assert rc == 0, (rc, L.mgp_error_string(rc))
    mgp.check(L.mgp_stream_sync(None))
    for _ in range(ncycles):
        s.cycle(cyc, nsmooth, omega)
    rel = s.relres()
    s.to_host()
    want = interior(np.array(s.array("u"), copy=True).reshape(u.pshape), ud)
    got = interior(u.download(), ud)
    np.testing.assert_array_equal(got, want)
    fnorm = np.abs(interior(np.array(s.array("f"), copy=True).reshape(u.pshape), ud)).max()
    norm = nrm.download()[0]
    say("  norm %.6e relres %.6e" % (norm / fnorm, rel))
    assert norm / fnorm == rel
    s.close()


def test_noalloc_cycles_equal_solver_cycles_twice_then_smaller():
    for case in CASES:
        run_case(case)


def test_noalloc_rejects_what_does_not_fit():
    L = mgp.lib()
    F = mgp.Faces()
    for q in range(6):
        F.type[q] = mgp.DIRICHLET
    F.h = 1.0 / 256
    big = mgp.DeviceArray((257, 257, 129))  # hierarchy above the scratch size
    assert L.mgp_cycle_noalloc_f64(b"Vcycle", big.ref(), big.ref(), ctypes.byref(F), 6, 1, 2, 1.0, 0, None,
                                   None) == mgp.ERR_ARG
    big.free()
    u = mgp.DeviceArray((33, 33, 33), (1, 0, 0, 0, 0, 0))  # halo on a Dirichlet face
    assert L.mgp_cycle_noalloc_f64(b"Vcycle", u.ref(), u.ref(), ctypes.byref(F), 3, 1, 2, 1.0, 0, None,
                                   None) == mgp.ERR_ARG
    v = mgp.DeviceArray((33, 33, 33))
    assert L.mgp_cycle_noalloc_f64(b"Wcycle", v.ref(), v.ref(), ctypes.byref(F), 3, 1, 2, 1.0, 0, None,
                                   None) == mgp.ERR_ARG
    assert L.mgp_cycle_noalloc_f64(b"Fcycle", v.ref(), v.ref(), ctypes.byref(F), 1, 1, 2, 1.0, 0, None,
                                   None) == mgp.ERR_ARG
    w = mgp.DeviceArray((34, 33, 33))  # does not coarsen
    assert L.mgp_cycle_noalloc_f64(b"Vcycle", w.ref(), w.ref(), ctypes.byref(F), 3, 1, 2, 1.0, 0, None,
                                   None) == mgp.ERR_ARG
