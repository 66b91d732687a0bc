"""Pins the oracle's hi + lo refinement iterate (og_hilo_update,
og_hilo_residual: FP64 hi plus FP32 lo, this project's addition for
mixed-precision refinement) against exact rational arithmetic.

The residual of u = hi + lo is pinned on inputs where every FP64 operation
the oracle performs is exact (hi multiples of 2^-6, lo multiples of 2^-40,
h = 1/8, face data multiples of 1/4), so the oracle must return exactly
f - Lap_h(hi + lo) with Neumann ghosts u(mirror) + sign*2h*g
(include/neumann.h:112-130, include/residual.h:58-79 of the reference),
computed here with fractions.  A dropped lo term, a wrong ghost sign,
mirror or face index each change r there.  The update is pinned by the
TwoSum identity on random inputs."""
import random
import zlib
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O

HILO_CASES = [((6, 5, 7), "DDDDDD"), ((6, 5, 7), "NDNDND"), ((5, 7, 6), "DNDNDN"), ((4, 4, 5), "NNNNND")]


def _face_shape(n, q):
    m = list(n)
    m[{O.N: 1, O.S: 1, O.E: 0, O.W: 0, O.T: 2, O.B: 2}[q]] = 1
    return tuple(m)


def _exact_residual(n, hi, lo, f, faces, box, h):
    """r = f - Lap_h(u) with u = hi + lo as fractions; the neighbour of a
    point beyond a Neumann face is u(mirror) + sign*2h*g."""
    hF = Fraction(h)
    u = {}
    for idx in np.ndindex(*n):
        u[idx] = Fraction(float(hi[idx])) + Fraction(float(lo[idx]))
    face_of = {(0, -1): O.W, (0, 1): O.E, (1, -1): O.S, (1, 1): O.N, (2, -1): O.B, (2, 1): O.T}
    out = {}
    for i in range(box[0], box[1]):
        for j in range(box[2], box[3]):
            for k in range(box[4], box[5]):
                p = (i, j, k)
                lap = Fraction(0)
                for ax in range(3):
                    for d in (-1, 1):
                        a = list(p)
                        a[ax] += d
                        if 0 <= a[ax] < n[ax]:
                            v = u[tuple(a)]
                        else:
                            q = face_of[(ax, d)]
                            a[ax] -= 2 * d
                            gi = list(p)
                            gi[ax] = 0
                            g = Fraction(float(faces[q].reshape(_face_shape(n, q))[tuple(gi)]))
                            v = u[tuple(a)] + d * 2 * hF * g
                        lap += v - u[p]
                out[p] = Fraction(float(f[p])) - lap / (hF * hF)
    return out


@pytest.mark.parametrize("shape,bc", HILO_CASES, ids=[str(i) for i in range(len(HILO_CASES))])
def test_hilo_residual_is_exact_on_representable_inputs(shape, bc):
    rng = np.random.default_rng(zlib.crc32(repr((shape, bc, "hilo-exact")).encode()))
    halo = [1 if c == "N" else 0 for c in bc]
    L = O.layout(shape, halo)
    size = L.size
    hi = rng.integers(-256, 257, size).astype(np.float64) / 64.0
    lo = (rng.integers(-(1 << 20), (1 << 20) + 1, size).astype(np.float64) * 2.0 ** -40).astype(np.float32)
    f = rng.integers(-64, 65, size).astype(np.float64) / 8.0
    faces = {q: rng.integers(-16, 17, int(np.prod(_face_shape(shape, q)))).astype(np.float64) / 4.0
             for q in range(6) if bc[q] == "N"}
    box = [int(bc[O.W] == "D"), shape[0] - int(bc[O.E] == "D"), int(bc[O.S] == "D"), shape[1] - int(bc[O.N] == "D"),
           int(bc[O.B] == "D"), shape[2] - int(bc[O.T] == "D")]
    r = np.full(size, 7.0)
    nrm = O.hilo_residual(L, hi, lo, f, r, faces, box, 0.125)
    want = _exact_residual(shape, O.interior(L, hi), O.interior(L, lo), O.interior(L, f), faces, box, 0.125)
    ri = O.interior(L, r)
    vmax = 0.0
    for p, v in want.items():
        assert Fraction(float(ri[p])) == v, (p, ri[p], float(v))
        vmax = max(vmax, abs(float(v)))
    # points outside the box keep their stored r (7.0) and count in the norm
    outside = sum(1 for p in np.ndindex(*shape) if p not in want)
    assert nrm == (max(vmax, 7.0) if outside else vmax)
    # the lo word matters: the same with lo = 0 differs where lo varies
    r2 = np.full(size, 7.0)
    O.hilo_residual(L, hi, np.zeros_like(lo), f, r2, faces, box, 0.125)
    assert not np.array_equal(r2, r)


def test_hilo_update_is_an_error_free_renormalisation():
    rng = np.random.default_rng(zlib.crc32(b"hilo-update"))
    shape = (9, 8, 7)
    L = O.layout(shape)
    n = L.size
    hi0 = rng.standard_normal(n) * 3.0
    lo0 = (rng.standard_normal(n) * 1e-17).astype(np.float32)
    e = (rng.standard_normal(n) * 10.0 ** rng.integers(-12, 1, n)).astype(np.float32)
    hi, lo = hi0.copy(), lo0.copy()
    O.hilo_update(L, hi, lo, e)
    t = e.astype(np.float64) + lo0.astype(np.float64)  # fl(e + lo)
    np.testing.assert_array_equal(hi, hi0 + t)        # hi' = fl(hi + t)
    random.seed(1)
    for m in random.sample(range(n), 300):
        exact = Fraction(float(hi0[m])) + Fraction(float(e[m])) + Fraction(float(lo0[m]))
        got = Fraction(float(hi[m])) + Fraction(float(lo[m]))
        # rounding of t to FP64 and of the TwoSum error to FP32 only
        bound = abs(Fraction(float(t[m]))) * Fraction(1, 2 ** 53) + abs(Fraction(float(lo[m]))) * Fraction(1, 2 ** 23)
        assert abs(got - exact) <= bound, (m, float(got - exact), float(bound))
        # lo' is at most half an ulp of hi'
        assert abs(float(lo[m])) <= np.spacing(abs(hi[m])) / 2 * (1 + 2.0 ** -20)
