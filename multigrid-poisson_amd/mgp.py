"""ctypes binding of the mgp C ABI (include/mgp.h, lib/libmgp.so).

The same entry points a C, Go (cgo) or Java (JNI) caller would bind.  There is
no fallback: if lib/libmgp.so is missing or no GPU is visible, calls raise.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MGP_LIBRARY: another build of the library (same C ABI), e.g. with other
# build-time cache policies, for measurements
LIB_PATH = os.environ.get("MGP_LIBRARY") or os.path.join(HERE, "lib", "libmgp.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mgp.h")

NORTH, SOUTH, EAST, WEST, TOP, BOTTOM = range(6)
DIRICHLET, NEUMANN = 0, 1
REFERENCE, CONSISTENT = 0, 1
INJECTION, FULL_WEIGHTING = 0, 1
JACOBI, GAUSS_SEIDEL = 0, 1
F64, F32, MIXED = 0, 1, 2
NO_GRAPHS, NO_FUSED, QUIET, PLAIN_ITERATE, COARSE_FP32, COARSE_CG = 1, 2, 4, 8, 16, 32
OK, ERR_ARG, ERR_NODEV, ERR_STATE = 0, -1, -2, -3
LOC = "NSEWTB"


class MgpError(RuntimeError):
    pass


class Array(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("shape", ctypes.c_int64 * 3), ("halo", ctypes.c_int64 * 6),
                ("stride", ctypes.c_int64 * 3)]


class Faces(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32 * 6), ("g", ctypes.c_void_p * 6), ("h", ctypes.c_double)]


class SolverConfig(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int64 * 3), ("levels", ctypes.c_int64), ("h", ctypes.c_double),
                ("origin", ctypes.c_double * 3), ("bc", ctypes.c_int32 * 6), ("dtype", ctypes.c_int32),
                ("restriction", ctypes.c_int32), ("relaxation", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("maxiter", ctypes.c_int64), ("miniter", ctypes.c_int64),
                ("tolerance", ctypes.c_double), ("coarse_sweeps", ctypes.c_int64), ("refine_polish", ctypes.c_int64),
                ("half_from", ctypes.c_int64)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int64), ("levels", ctypes.c_int64), ("relaxation", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("nsmooth", ctypes.c_int64), ("omega", ctypes.c_double),
                ("coarse_sweeps", ctypes.c_int64), ("maxiter", ctypes.c_int64), ("miniter", ctypes.c_int64),
                ("tolerance", ctypes.c_double)]


BATCH_NO_GRAPHS, BATCH_RETUNE = 1, 2
MAX_MODES = 16


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("max_active", ctypes.c_int32), ("relaxation", ctypes.c_int32),
                ("nsmooth", ctypes.c_int64), ("omega", ctypes.c_double), ("coarse_sweeps", ctypes.c_int64),
                ("maxiter", ctypes.c_int64), ("tolerance", ctypes.c_double), ("flags", ctypes.c_int32)]


SERVER_GRAPHS = 1
SERVER_DIRECT_ALLOC = 2


class Request(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int64 * 3), ("bc", ctypes.c_int32 * 6), ("nmodes", ctypes.c_int32),
                ("abc", ctypes.c_int32 * (3 * MAX_MODES)), ("amp", ctypes.c_double * MAX_MODES),
                ("u_out", ctypes.c_void_p)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("iterations", ctypes.c_int64), ("rel_res", ctypes.c_double),
                ("latency_s", ctypes.c_double), ("queue_s", ctypes.c_double), ("run_s", ctypes.c_double),
                ("setup_s", ctypes.c_double), ("fetch_s", ctypes.c_double)]

_lib = None
P = ctypes.POINTER
vp, i32, i64, d, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_float
BOX = P(ctypes.c_int64 * 6)

_SIGS = {
    "mgp_device_count": [P(ctypes.c_int)], "mgp_set_device": [i32], "mgp_malloc": [P(vp), ctypes.c_size_t],
    "mgp_free": [vp], "mgp_host_alloc": [P(vp), ctypes.c_size_t], "mgp_host_free": [vp],
    "mgp_memcpy_h2d": [vp, vp, ctypes.c_size_t, vp], "mgp_memcpy_d2h": [vp, vp, ctypes.c_size_t, vp],
    "mgp_memcpy_d2d": [vp, vp, ctypes.c_size_t, vp], "mgp_stream_create": [P(vp)], "mgp_stream_destroy": [vp],
    "mgp_stream_sync": [vp], "mgp_mem_info": [P(ctypes.c_size_t), P(ctypes.c_size_t)],
    "mgp_zero_scalars": [vp, i64, vp],
    "mgp_cast_f64_f32": [P(Array), P(Array), vp], "mgp_cast_f32_f64": [P(Array), P(Array), vp],
    "mgp_add_f32_to_f64": [P(Array), P(Array), vp],
    "mgp_update_residual_f64_f32": [i32, P(Array), P(Array), P(Array), P(Array), P(Array), P(Array), P(Faces), BOX,
                                    vp, vp],
    "mgp_update_residual_hilo_f64_f32": [P(Array), P(Array), P(Array), P(Array), P(Array), P(Array), P(Array),
                                         P(Array), P(Faces), BOX, vp, vp],
    "mgp_residual_hilo_f64": [P(Array), P(Array), P(Array), P(Array), P(Faces), BOX, vp, vp],
    "mgp_graph_begin": [vp], "mgp_graph_end": [vp, P(vp)], "mgp_graph_launch": [vp, vp], "mgp_graph_destroy": [vp],
    "mgp_event_create": [P(vp)], "mgp_event_destroy": [vp], "mgp_event_record": [vp, vp], "mgp_event_sync": [vp],
    "mgp_event_elapsed_ms": [vp, vp, P(f)],
    "mgp_tuning_set": [ctypes.c_char_p, i64], "mgp_tuning_get": [ctypes.c_char_p, P(i64)],
    "mgp_solver_create": [P(SolverConfig), P(vp)], "mgp_solver_destroy": [vp],
    "mgp_solver_init_problem": [vp, i32], "mgp_solver_init_zero": [vp], "mgp_solver_to_device": [vp], "mgp_solver_to_device_zero": [vp],
    "mgp_solver_to_host": [vp], "mgp_solver_solve": [vp, ctypes.c_char_p, i64, d, d],
    "mgp_solver_cycle": [vp, ctypes.c_char_p, i64, d], "mgp_solver_relres": [vp, P(d)], "mgp_solver_sync": [vp], "mgp_solver_reset": [vp],
    "mgp_solver_stats": [vp, P(d), P(d), P(i64)], "mgp_solver_array": [vp, i32, i32, P(vp), P(Array)],
    "mgp_solver_face": [vp, i32, i32, P(vp), P(Array), P(ctypes.c_int)], "mgp_solver_save": [vp, ctypes.c_char_p],
    "mgp_fill_modes_f64": [P(Array), i64, P(ctypes.c_int32), P(d), d, vp], "mgp_pack_f64": [P(Array), vp, vp],
    "mgp_residual_cast_f64_f32": [i32, P(Array), P(Array), P(Array), P(Array), BOX, d, vp, vp],
    "mgp_batch_create": [P(BatchConfig), P(i64), P(vp)], "mgp_batch_destroy": [vp],
    "mgp_batch_set_rhs": [vp, i64, P(d)], "mgp_batch_solve": [vp, P(d)], "mgp_batch_stats": [vp, i64, P(d), P(i64)],
    "mgp_batch_solution": [vp, i64, P(d)], "mgp_batch_tile": [vp, i64, i64, P(i64)],
    "mgp_server_create": [P(ServerConfig), P(vp)], "mgp_server_destroy": [vp],
    "mgp_server_submit": [vp, P(Request), P(i64)], "mgp_server_wait": [vp, i64, P(Result)],
    "mgp_server_stats": [vp, P(i64), P(i64), P(i64)],
    "mgp_fill_f16": [P(Array), d, vp], "mgp_rbgs_half_f16": [P(Array), P(Array), BOX, d, d, i32, vp],
    "mgp_rbgs_sweeps_f16": [P(Array), P(Array), BOX, d, d, i64, vp],
    "mgp_rbgs_sweeps_f32_f64": [i32, P(Array), P(Array), P(Array), P(Faces), BOX, d, i64, vp],
    "mgp_cg_solve_f64": [P(Array), P(Array), P(Faces), BOX, i64, d, vp],
    "mgp_cg_solve_f32": [P(Array), P(Array), P(Faces), BOX, i64, d, vp],
    "mgp_cycle_noalloc_f64": [ctypes.c_char_p, P(Array), P(Array), P(Faces), i64, i64, i64, d, i64, vp, vp],
    "mgp_residual_f16": [P(Array), P(Array), P(Array), BOX, d, vp, vp],
    "mgp_restrict_fw_scaled_f16": [P(Array), P(Array), vp, vp],
    "mgp_restrict_fw_scaled_f32_f16": [P(Array), P(Array), vp, vp],
    "mgp_prolong_add_scaled_f16": [P(Array), P(Array), vp, vp],
    "mgp_prolong_add_scaled_f16_f32": [P(Array), P(Array), vp, vp],
}
for sfx in ("f64", "f32"):
    _SIGS.update({
        "mgp_fill_" + sfx: [P(Array), d, vp], "mgp_add_" + sfx: [P(Array), P(Array), vp],
        "mgp_norm_inf_" + sfx: [P(Array), vp, vp],
        "mgp_rbgs_half_" + sfx: [i32, P(Array), P(Array), BOX, d, d, i32, vp],
        "mgp_rbgs_sweep_" + sfx: [i32, P(Array), P(Array), P(Array), P(Faces), BOX, d, vp],
        "mgp_rbgs_sweep2_" + sfx: [i32, P(Array), P(Array), P(Array), P(Faces), BOX, d, vp],
        "mgp_rbgs_sweep2_prolong_" + sfx: [i32, P(Array), P(Array), P(Array), P(Faces), BOX, d, P(Array), vp],
        "mgp_rbgs_sweep2_from_zero_" + sfx: [i32, P(Array), P(Array), P(Faces), BOX, d, vp],
        "mgp_rbgs_sweeps_" + sfx: [i32, P(Array), P(Array), P(Array), P(Faces), BOX, d, i64, vp],
        "mgp_residual_restrict_fw_" + sfx: [i32, P(Array), P(Array), P(Array), P(Array), BOX, d, vp],
        "mgp_residual_norm_" + sfx: [i32, P(Array), P(Array), P(Array), BOX, d, vp, vp],
        "mgp_jacobi_" + sfx: [i32, P(Array), P(Array), P(Array), BOX, d, d, vp],
        "mgp_residual_" + sfx: [i32, P(Array), P(Array), P(Array), BOX, d, vp, vp],
        "mgp_restrict_inject_" + sfx: [P(Array), P(Array), vp], "mgp_restrict_fw_" + sfx: [P(Array), P(Array), vp],
        "mgp_prolong_trilinear_" + sfx: [P(Array), P(Array), vp], "mgp_prolong_add_" + sfx: [P(Array), P(Array), vp],
        "mgp_bc_dirichlet_" + sfx: [P(Array), i32, vp, vp], "mgp_bc_neumann_" + sfx: [P(Array), i32, vp, d, vp],
        "mgp_bc_neumann_faces_" + sfx: [P(Array), P(Faces), vp],
        "mgp_bc_restrict_neumann_" + sfx: [i32, P(Array), i32, vp, vp, d, vp],
    })


def lib():
    """Load lib/libmgp.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MgpError("HIP library not built: %s (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.mgp_error_string.argtypes = [ctypes.c_int]
        L.mgp_error_string.restype = ctypes.c_char_p
        L.mgp_version.restype = ctypes.c_char_p
        L.mgp_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def declared_symbols(header=HEADER):
    """Every function name declared in include/mgp.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(mgp_\w+)\s*\(", text, flags=re.M)))


def check(rc, what=""):
    if rc != 0:
        msg = lib().mgp_error_string(rc).decode()
        extra = lib().mgp_last_error().decode()
        raise MgpError("%s failed: %s (%d)%s" % (what, msg, rc, (": " + extra) if extra else ""))


def device_count():
    n = ctypes.c_int(0)
    lib().mgp_device_count(ctypes.byref(n))
    return n.value


def require_gpu():
    if device_count() < 1:
        raise MgpError("no GPU visible to the HIP runtime")


def box(b):
    return ctypes.byref((ctypes.c_int64 * 6)(*[int(x) for x in b]))


def padded_shape(shape, halo):
    hN, hS, hE, hW, hT, hB = halo
    return (shape[0] + hE + hW, shape[1] + hN + hS, shape[2] + hT + hB)


class DeviceArray:
    """A halo-padded device array (own allocation) for kernel-level calls."""

    def __init__(self, shape, halo=(0, 0, 0, 0, 0, 0), dtype=np.float64):
        self.shape = tuple(int(x) for x in shape)
        self.halo = tuple(int(x) for x in halo)
        self.dtype = np.dtype(dtype)
        self.pshape = padded_shape(self.shape, self.halo)
        self.nbytes = int(np.prod(self.pshape)) * self.dtype.itemsize
        p = vp()
        check(lib().mgp_malloc(ctypes.byref(p), self.nbytes), "mgp_malloc")
        self.ptr = p.value
        a = Array()
        a.data = self.ptr
        a.shape[:] = list(self.shape)
        a.halo[:] = list(self.halo)
        s1 = self.pshape[2]
        a.stride[:] = [self.pshape[1] * s1, s1, 1]
        self.desc = a

    @property
    def sfx(self):
        return {8: "f64", 4: "f32", 2: "f16"}[self.dtype.itemsize]

    def upload(self, host):
        h = np.ascontiguousarray(host, dtype=self.dtype).reshape(self.pshape)
        check(lib().mgp_memcpy_h2d(self.ptr, h.ctypes.data, self.nbytes, None), "h2d")
        check(lib().mgp_stream_sync(None), "sync")
        return self

    def download(self):
        h = np.empty(self.pshape, dtype=self.dtype)
        check(lib().mgp_stream_sync(None), "sync")
        check(lib().mgp_memcpy_d2h(h.ctypes.data, self.ptr, self.nbytes, None), "d2h")
        check(lib().mgp_stream_sync(None), "sync")
        return h

    def ref(self):
        return ctypes.byref(self.desc)

    def free(self):
        if getattr(self, "ptr", None):
            lib().mgp_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBuffer:
    """Raw device memory (face arrays, scalars)."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        p = vp()
        check(lib().mgp_malloc(ctypes.byref(p), max(self.nbytes, 8)), "mgp_malloc")
        self.ptr = p.value
        check(lib().mgp_memcpy_h2d(self.ptr, self.host.ctypes.data, self.nbytes, None), "h2d")
        check(lib().mgp_stream_sync(None), "sync")

    def download(self):
        out = np.empty_like(self.host)
        check(lib().mgp_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, None), "d2h")
        check(lib().mgp_stream_sync(None), "sync")
        return out

    def __del__(self):
        try:
            if self.ptr:
                lib().mgp_free(self.ptr)
        except Exception:
            pass


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


class Solver:
    """PoissonSolver<T, R, TrilinearInterpolation, S> through the C ABI."""

    def __init__(self, dims, levels, h, origin=(0.0, 0.0, 0.0), bc="DDDDDD", mode=REFERENCE,
                 relaxation=GAUSS_SEIDEL, restriction=INJECTION, problem=0, dtype=F64, flags=QUIET,
                 maxiter=100, miniter=1, tolerance=1e-6, coarse_sweeps=0, refine_polish=0,
                 half_from=0):
        require_gpu()
        c = SolverConfig()
        c.dims[:] = [int(x) for x in dims]
        c.levels = int(levels)
        c.h = float(h)
        c.origin[:] = [float(x) for x in origin]
        c.bc[:] = [NEUMANN if ch == "N" else DIRICHLET for ch in bc]
        c.dtype, c.restriction, c.relaxation = int(dtype), int(restriction), int(relaxation)
        c.mode, c.flags = int(mode), int(flags)
        c.maxiter, c.miniter, c.tolerance = int(maxiter), int(miniter), float(tolerance)
        c.coarse_sweeps = int(coarse_sweeps)
        c.refine_polish = int(refine_polish)
        c.half_from = int(half_from)
        self.config = c
        self.levels = int(levels)
        self.np_dtype = np.float32 if dtype == F32 else np.float64
        h_ = vp()
        check(lib().mgp_solver_create(ctypes.byref(c), ctypes.byref(h_)), "mgp_solver_create")
        self.h = h_.value
        if problem is None:
            pass
        elif problem < 0:
            check(lib().mgp_solver_init_zero(self.h), "init_zero")
        else:
            check(lib().mgp_solver_init_problem(self.h, int(problem)), "init")

    def close(self):
        if getattr(self, "h", None):
            lib().mgp_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, level, which):
        host = vp()
        a = Array()
        check(lib().mgp_solver_array(self.h, int(level), ord(which), ctypes.byref(host), ctypes.byref(a)), "array")
        return host.value, a

    def layout(self, level=0):
        return self._arr(level, "u")[1]

    def lo_word(self):
        """FP32 lo word of a mixed solver's hi + lo iterate (level 0, flat,
        halos included); the iterate is array("u") + lo_word()."""
        _, a = self._arr(0, "l")
        n = a.stride[0] * (a.shape[0] + a.halo[2] + a.halo[3])
        out = np.empty(n, np.float32)
        self.sync()
        check(lib().mgp_memcpy_d2h(out.ctypes.data, a.data, out.nbytes, None), "d2h")
        check(lib().mgp_stream_sync(None), "sync")
        return out

    def array(self, which, level=0):
        """Flat numpy view (shares memory) of the host mirror of u/p/f/r."""
        host, a = self._arr(level, which)
        pshape = padded_shape(list(a.shape), list(a.halo))
        n = int(np.prod(pshape))
        ct = ctypes.c_double if self.np_dtype == np.float64 else ctypes.c_float
        return np.ctypeslib.as_array(ctypes.cast(host, P(ct)), shape=(n,))

    def device_array(self, which, level=0):
        return self._arr(level, which)[1]

    def face(self, loc, level=0):
        host = vp()
        a = Array()
        t = ctypes.c_int()
        check(lib().mgp_solver_face(self.h, int(level), int(loc), ctypes.byref(host), ctypes.byref(a),
                                    ctypes.byref(t)), "face")
        n = int(np.prod(list(a.shape)))
        ct = ctypes.c_double if self.np_dtype == np.float64 else ctypes.c_float
        return np.ctypeslib.as_array(ctypes.cast(host.value, P(ct)), shape=(n,))

    def face_type(self, loc, level=0):
        host = vp()
        a = Array()
        t = ctypes.c_int()
        check(lib().mgp_solver_face(self.h, int(level), int(loc), ctypes.byref(host), ctypes.byref(a),
                                    ctypes.byref(t)), "face")
        return t.value

    def to_device(self):
        check(lib().mgp_solver_to_device(self.h), "to_device")

    def to_host(self):
        check(lib().mgp_solver_to_host(self.h), "to_host")

    def solve(self, cycle="Vcycle", nsmooth=10, omega=-1.0, maxtime=1200.0):
        check(lib().mgp_solver_solve(self.h, cycle.encode(), int(nsmooth), float(omega), float(maxtime)), "solve")
        return self.stats()

    def cycle(self, cycle="Vcycle", nsmooth=10, omega=-1.0):
        check(lib().mgp_solver_cycle(self.h, cycle.encode(), int(nsmooth), float(omega)), "cycle")

    def relres(self):
        r = ctypes.c_double()
        check(lib().mgp_solver_relres(self.h, ctypes.byref(r)), "relres")
        return r.value

    def sync(self):
        check(lib().mgp_solver_sync(self.h), "sync")

    def reset(self):
        check(lib().mgp_solver_reset(self.h), "reset")

    def stats(self):
        rr, t, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        check(lib().mgp_solver_stats(self.h, ctypes.byref(rr), ctypes.byref(t), ctypes.byref(it)), "stats")
        return it.value, rr.value, t.value

    def save(self, path):
        check(lib().mgp_solver_save(self.h, path.encode()), "save")


class Batch:
    """Batched FP32-cycle / FP64-refinement solves of zero-Dirichlet problems
    (mgp_batch_*): dims is a list of (nx, ny, nz)."""

    def __init__(self, dims, relaxation=JACOBI, nsmooth=3, omega=-1.0, levels=0, coarse_sweeps=0, maxiter=50,
                 miniter=1, tolerance=1e-10, flags=0):
        require_gpu()
        self.dims = [tuple(int(v) for v in d) for d in dims]
        c = BatchConfig()
        c.count, c.levels, c.relaxation, c.flags = len(self.dims), int(levels), int(relaxation), int(flags)
        c.nsmooth, c.omega, c.coarse_sweeps = int(nsmooth), float(omega), int(coarse_sweeps)
        c.maxiter, c.miniter, c.tolerance = int(maxiter), int(miniter), float(tolerance)
        flat = (ctypes.c_int64 * (3 * len(self.dims)))(*[v for d in self.dims for v in d])
        h = vp()
        check(lib().mgp_batch_create(ctypes.byref(c), flat, ctypes.byref(h)), "mgp_batch_create")
        self.h = h.value

    def close(self):
        if getattr(self, "h", None):
            lib().mgp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rhs(self, p, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.size == int(np.prod(self.dims[p]))
        check(lib().mgp_batch_set_rhs(self.h, int(p), f.ctypes.data_as(P(d))), "mgp_batch_set_rhs")

    def solve(self):
        t = ctypes.c_double()
        check(lib().mgp_batch_solve(self.h, ctypes.byref(t)), "mgp_batch_solve")
        return t.value

    def stats(self, p):
        rr, it = ctypes.c_double(), ctypes.c_int64()
        check(lib().mgp_batch_stats(self.h, int(p), ctypes.byref(rr), ctypes.byref(it)), "mgp_batch_stats")
        return it.value, rr.value

    def solution(self, p):
        out = np.empty(self.dims[p], dtype=np.float64)
        check(lib().mgp_batch_solution(self.h, int(p), out.ctypes.data_as(P(d))), "mgp_batch_solution")
        return out

    def tile(self, p, level):
        out = (ctypes.c_int64 * 2)()
        check(lib().mgp_batch_tile(self.h, int(p), int(level), out), "mgp_batch_tile")
        return out[0], out[1]


def make_request(dims, bc, modes, u_out=None):
    """mgp_request for a zero-boundary-data solve with f = sum of Fourier modes;
    modes = [(a, b, c, amp), ...]; bc = 6 characters 'D'/'N' (N S E W T B)."""
    r = Request()
    r.dims[:] = [int(v) for v in dims]
    r.bc[:] = [NEUMANN if ch == "N" else DIRICHLET for ch in bc]
    r.nmodes = len(modes)
    for m, (a, b, c, amp) in enumerate(modes):
        r.abc[3 * m], r.abc[3 * m + 1], r.abc[3 * m + 2] = int(a), int(b), int(c)
        r.amp[m] = float(amp)
    r.u_out = u_out.ctypes.data if u_out is not None else None
    return r


class Server:
    """The serving loop (mgp_server_*)."""

    def __init__(self, hbm_budget=64 << 30, max_active=4, relaxation=GAUSS_SEIDEL, nsmooth=2, omega=1.0,
                 coarse_sweeps=32, maxiter=100, tolerance=1e-8, flags=0):
        require_gpu()
        c = ServerConfig()
        c.hbm_budget, c.max_active, c.relaxation = int(hbm_budget), int(max_active), int(relaxation)
        c.nsmooth, c.omega, c.coarse_sweeps = int(nsmooth), float(omega), int(coarse_sweeps)
        c.maxiter, c.tolerance, c.flags = int(maxiter), float(tolerance), int(flags)
        h = vp()
        check(lib().mgp_server_create(ctypes.byref(c), ctypes.byref(h)), "mgp_server_create")
        self.h = h.value
        self._keep = {}

    def submit(self, dims, bc, modes, want_solution=True):
        """Queue a request; returns its id (the solution array is kept until wait)."""
        out = np.empty(tuple(int(v) for v in dims), dtype=np.float64) if want_solution else None
        r = make_request(dims, bc, modes, out)
        i = ctypes.c_int64()
        check(lib().mgp_server_submit(self.h, ctypes.byref(r), ctypes.byref(i)), "mgp_server_submit")
        self._keep[i.value] = out
        return i.value

    def wait(self, rid):
        res = Result()
        check(lib().mgp_server_wait(self.h, int(rid), ctypes.byref(res)), "mgp_server_wait")
        return res, self._keep.pop(rid)

    def stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().mgp_server_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "mgp_server_stats")
        return a.value, b.value, c.value

    def close(self):
        if getattr(self, "h", None):
            lib().mgp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
