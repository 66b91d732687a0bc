"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker, never the thing measured: only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

N, S, E, W, T, B = range(6)
DIRICHLET, NEUMANN = 0, 1
REFERENCE, CONSISTENT = 0, 1
JACOBI, GAUSS_SEIDEL = 0, 1
INJECTION, FULL_WEIGHTING = 0, 1
RELAXATION, VCYCLE, FCYCLE = 0, 1, 2
LOC_NAMES = "NSEWTB"


class Layout(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64 * 3), ("halo", ctypes.c_int64 * 6),
                ("s", ctypes.c_int64 * 3), ("size", ctypes.c_int64)]


class Params(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int64 * 3), ("levels", ctypes.c_int64), ("h", ctypes.c_double),
                ("origin", ctypes.c_double * 3), ("bc", ctypes.c_int32 * 6), ("mode", ctypes.c_int32),
                ("relaxation", ctypes.c_int32), ("restriction", ctypes.c_int32),
                ("problem", ctypes.c_int32), ("threads", ctypes.c_int32), ("coarse_sweeps", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        d, i64, i32, vp = ctypes.c_double, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
        L.og_layout_make.argtypes = [P(Layout), i64, i64, i64, P(i64 * 6)]
        L.og_residual.argtypes = [i32, P(Layout), vp, vp, vp, P(i64 * 6), d]
        L.og_gs_half.argtypes = [i32, P(Layout), vp, vp, P(i64 * 6), d, d, i32]
        L.og_jacobi.argtypes = [i32, P(Layout), vp, vp, vp, P(i64 * 6), d, d]
        for fn in ("og_restrict_inject", "og_restrict_fw", "og_prolong_trilinear"):
            getattr(L, fn).argtypes = [P(Layout), vp, P(Layout), vp]
        L.og_bc_dirichlet.argtypes = [P(Layout), vp, i32, vp]
        L.og_bc_neumann.argtypes = [P(Layout), vp, i32, vp, d]
        L.og_bc_restrict_neumann.argtypes = [i32, P(Layout), vp, i32, vp, vp, d]
        L.og_norm_inf.argtypes = [P(Layout), vp]
        L.og_norm_inf.restype = d
        for fn in ("og_ufun", "og_ffun", "og_dudx", "og_dudy", "og_dudz"):
            getattr(L, fn).argtypes = [i32, d, d, d]
            getattr(L, fn).restype = d
        L.og_create.argtypes = [P(Params)]
        L.og_create.restype = vp
        L.og_destroy.argtypes = [vp]
        L.og_to_device.argtypes = [vp]
        L.og_cycle.argtypes = [vp, i32, i64, d]
        L.og_solve.argtypes = [vp, i32, i64, d, i64, i64, d, vp, P(d)]
        L.og_solve.restype = i64
        L.og_relax_level.argtypes = [vp, i32, d, i64]
        L.og_residual_level.argtypes = [vp, i32]
        L.og_relres.argtypes = [vp]
        L.og_relres.restype = d
        L.og_array.argtypes = [vp, i32, i32]
        L.og_array.restype = P(d)
        L.og_array_layout.argtypes = [vp, i32]
        L.og_array_layout.restype = P(Layout)
        L.og_face.argtypes = [vp, i32, i32]
        L.og_face.restype = P(d)
        L.og_face_layout.argtypes = [vp, i32, i32]
        L.og_face_layout.restype = P(Layout)
        L.og_face_type.argtypes = [vp, i32, i32]
        L.og_face_type.restype = ctypes.c_int32
        L.og_level_h.argtypes = [vp, i32]
        L.og_level_h.restype = d
        L.og_f2h_n.argtypes = [vp, vp, i64]
        L.og_h2f_n.argtypes = [vp, vp, i64]
        L.og_h_scale.argtypes = [d, i32]
        L.og_h_scale.restype = ctypes.c_float
        L.og_h_gs_half.argtypes = [P(Layout), vp, vp, P(i64 * 6), d, d, i32]
        L.og_h_residual.argtypes = [P(Layout), vp, vp, vp, P(i64 * 6), d]
        L.og_h_residual.restype = d
        L.og_h_restrict_fw.argtypes = [P(Layout), vp, i32, P(Layout), vp, d]
        L.og_h_prolong_add.argtypes = [P(Layout), vp, P(Layout), vp, i32, d]
        L.og_hilo_update.argtypes = [P(Layout), vp, vp, vp]
        L.og_hilo_residual.argtypes = [P(Layout), vp, vp, vp, vp, P(vp * 6), P(i64 * 6), d]
        L.og_hilo_residual.restype = d
        _lib = L
    return _lib


def layout(n, halo=(0, 0, 0, 0, 0, 0)):
    """Halo-padded layout of an (nx, ny, nz) array; halo in N,S,E,W,T,B order."""
    L = Layout()
    lib().og_layout_make(ctypes.byref(L), int(n[0]), int(n[1]), int(n[2]),
                         ctypes.byref((ctypes.c_int64 * 6)(*halo)))
    return L


def padded_shape(L):
    h = list(L.halo)
    return (L.n[0] + h[E] + h[W], L.n[1] + h[N] + h[S], L.n[2] + h[T] + h[B])


def interior(L, a):
    """View of the shape region of a flat padded buffer, indexed [i, j, k]."""
    h = list(L.halo)
    a3 = a.reshape(padded_shape(L))
    return a3[h[W]:h[W] + L.n[0], h[S]:h[S] + L.n[1], h[B]:h[B] + L.n[2]]


def ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def box_arr(box):
    return (ctypes.c_int64 * 6)(*box)


class Solver:
    """One oracle solver: PoissonSolver<double, R, TrilinearInterpolation, S>."""

    def __init__(self, dims, levels, h, origin=(0.0, 0.0, 0.0), bc="DDDDDD", mode=REFERENCE,
                 relaxation=GAUSS_SEIDEL, restriction=INJECTION, problem=0, threads=0, coarse_sweeps=0):
        p = Params()
        p.dims[:] = [int(x) for x in dims]
        p.levels = int(levels)
        p.h = float(h)
        p.origin[:] = [float(x) for x in origin]
        p.bc[:] = [NEUMANN if c == "N" else DIRICHLET for c in bc]
        p.mode, p.relaxation, p.restriction = int(mode), int(relaxation), int(restriction)
        p.problem, p.threads, p.coarse_sweeps = int(problem), int(threads), int(coarse_sweeps)
        self.params = p
        self.levels = int(levels)
        self.h = lib().og_create(ctypes.byref(p))
        if not self.h:
            raise ValueError("invalid grid for %d levels: %s" % (levels, dims))

    def close(self):
        if getattr(self, "h", None):
            lib().og_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def layout(self, level=0):
        return lib().og_array_layout(self.h, level).contents

    def array(self, which, level=0):
        """Flat padded numpy view (shares memory) of u/f/r/p."""
        L = self.layout(level)
        p = lib().og_array(self.h, level, ord(which))
        return np.ctypeslib.as_array(p, shape=(L.size,))

    def face(self, loc, level=0):
        G = lib().og_face_layout(self.h, level, loc).contents
        p = lib().og_face(self.h, level, loc)
        return np.ctypeslib.as_array(p, shape=(G.size,))

    def face_type(self, loc, level=0):
        return lib().og_face_type(self.h, level, loc)

    def level_h(self, level=0):
        return lib().og_level_h(self.h, level)

    def to_device(self):
        lib().og_to_device(self.h)

    def cycle(self, cycle, nsmooth, omega=-1.0):
        lib().og_cycle(self.h, int(cycle), int(nsmooth), float(omega))

    def solve(self, cycle, nsmooth, omega=-1.0, maxiter=100, miniter=1, tol=1e-6):
        out = np.zeros(max(int(maxiter), 1))
        fin = ctypes.c_double()
        it = lib().og_solve(self.h, int(cycle), int(nsmooth), float(omega), int(maxiter), int(miniter),
                            float(tol), out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(fin))
        nrec = min(it + 1, int(maxiter))
        return int(it), fin.value, out[:nrec].copy()

    def relax(self, level, omega, nsmooth):
        lib().og_relax_level(self.h, int(level), float(omega), int(nsmooth))

    def residual(self, level=0):
        lib().og_residual_level(self.h, int(level))

    def relres(self):
        return lib().og_relres(self.h)


def cycle_id(name):
    n = name.lower()
    return {"vcycle": VCYCLE, "fcycle": FCYCLE, "relaxation": RELAXATION}[n]


# ---- FP16 storage, FP32 arithmetic (arrays without halos: numpy float16,
# or float32 where a function takes FP32 input or output)

def _hp(a, dtypes=(np.float16,)):
    assert a.dtype in dtypes and a.flags.c_contiguous, a.dtype
    return a.ctypes.data_as(ctypes.c_void_p)


def f2h(x):
    """float32 -> float16 by the oracle's own rounding (nearest, ties to even)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().og_f2h_n(x.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), x.size)
    return out.view(np.float16)


def h2f(x):
    x = np.ascontiguousarray(x, dtype=np.float16)
    out = np.empty(x.shape, dtype=np.float32)
    lib().og_h2f_n(x.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), x.size)
    return out


def h_scale(norm, inverse=False):
    return float(lib().og_h_scale(float(norm), int(inverse)))


def h_gs_half(u, f, box, omega, h, colour):
    """In place on float16 u."""
    lib().og_h_gs_half(ctypes.byref(layout(u.shape)), _hp(u), _hp(f), ctypes.byref(box_arr(box)), float(omega),
                       float(h), int(colour))


def h_residual(u, f, r, box, h):
    """r (float16) = f - L u on the box; returns max |r| over r's shape."""
    return lib().og_h_residual(ctypes.byref(layout(u.shape)), _hp(u), _hp(f), _hp(r), ctypes.byref(box_arr(box)),
                               float(h))


def h_restrict_fw(fine, coarse, norm):
    """coarse (float16) = sigma(norm) FW(fine), fine float16 or float32."""
    lib().og_h_restrict_fw(ctypes.byref(layout(fine.shape)), _hp(fine, (np.float16, np.float32)),
                           int(fine.dtype == np.float32), ctypes.byref(layout(coarse.shape)), _hp(coarse),
                           float(norm))


def h_prolong_add(coarse, fine, norm):
    """fine (float16 or float32) += P(coarse) / sigma(norm), in place."""
    lib().og_h_prolong_add(ctypes.byref(layout(coarse.shape)), _hp(coarse), ctypes.byref(layout(fine.shape)),
                           _hp(fine, (np.float16, np.float32)), int(fine.dtype == np.float32), float(norm))


# ---- the hi + lo iterate of mixed-precision refinement (FP64 hi, FP32 lo)

def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def hilo_update(L, hi, lo, e):
    """In place: hi, lo (flat padded float64 / float32) += e (float32)."""
    lib().og_hilo_update(ctypes.byref(L), ptr(hi), _fp(lo), _fp(e))


def hilo_residual(L, hi, lo, f, r, faces, box, h):
    """r = f - L(hi + lo) on the box; faces: dict face -> float64 face array
    for Neumann faces.  Returns max |r| over the box and r outside it."""
    g = (ctypes.c_void_p * 6)(*[ptr(faces[q]) if q in faces else None for q in range(6)])
    return lib().og_hilo_residual(ctypes.byref(L), ptr(hi), _fp(lo), ptr(f), ptr(r), ctypes.byref(g),
                                  ctypes.byref(box_arr(box)), float(h))
