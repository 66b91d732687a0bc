"""Pins the oracle's FP16-storage operators (oracle/mg_oracle.c, og_h_* and
og_f2h/og_h2f), which check the FP16 coarse-level kernels.

- Conversions against numpy's own binary16 conversion: every binary16 value,
  every midpoint between neighbours (the ties) and the floats either side of
  them, subnormals, overflow, and random float32 bit patterns.
- Each operator against the FP64 oracle operator it restates (og_gs_half,
  og_residual, og_restrict_fw, og_prolong_trilinear) on inputs where every
  float operation is exact: multiples of 1/8 with omega = 3/4 (omega/6 = 1/8)
  and h = 1/2, so the two must agree exactly once rounded to binary16; a
  wrong neighbour, sign, colour, weight or scale fails these.
- The same operators on random inputs within the binary16 rounding bound
  (half an ulp plus float rounding), which a truncating conversion fails.
"""
import numpy as np
import pytest

import oracle_lib as ol


# ------------------------------------------------------------ conversions

def test_h2f_every_binary16_value():
    bits = np.arange(65536, dtype=np.uint16)
    h = bits.view(np.float16)
    got = ol.h2f(h)
    want = h.astype(np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))  # signed zeros too


def _f2h_bits(x):
    return ol.f2h(x).view(np.uint16)


def test_f2h_round_trip_ties_and_neighbours():
    bits = np.arange(65536, dtype=np.uint16)
    h = bits.view(np.float16)
    fin = np.isfinite(h)
    vals = h[fin].astype(np.float32)
    assert np.array_equal(_f2h_bits(vals), bits[fin])  # every finite value converts back to itself
    # midpoints between neighbouring positive values (exact in float32), the
    # float32 values either side of them, and the same negated
    pos = np.sort(h[fin & (h > 0)].astype(np.float64))
    pos = np.concatenate([[0.0], pos, [65536.0]])  # 65536: the next step above the largest finite value
    mids = ((pos[:-1] + pos[1:]) / 2).astype(np.float32)
    assert np.array_equal(mids.astype(np.float64), (pos[:-1] + pos[1:]) / 2)
    probes = np.concatenate([mids, np.nextafter(mids, np.float32(np.inf)), np.nextafter(mids, np.float32(0))])
    probes = np.concatenate([probes, -probes])
    with np.errstate(over="ignore"):
        want = probes.astype(np.float16)
    assert np.array_equal(_f2h_bits(probes), want.view(np.uint16))


def test_f2h_random_float32_patterns_and_specials():
    rng = np.random.default_rng(16)
    x = rng.integers(0, 2 ** 32, size=2_000_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    # concentrate half of them where binary16 is finite or subnormal
    y = (rng.standard_normal(1_000_000) * np.exp2(rng.uniform(-30, 17, 1_000_000))).astype(np.float32)
    x = np.concatenate([x, y, np.array([0.0, -0.0, np.inf, -np.inf, 65504, 65519.99, 65520, 2.0 ** -24,
                                        2.0 ** -25, 1.5 * 2.0 ** -25, 2.0 ** -26, 1e-45], dtype=np.float32)])
    with np.errstate(over="ignore"):
        want = x.astype(np.float16)
    got = ol.f2h(x)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint16), want[~nan].view(np.uint16))


def test_scale_is_a_power_of_two_bringing_the_norm_to_half_one():
    rng = np.random.default_rng(3)
    for n in np.concatenate([np.exp2(rng.uniform(-100, 100, 500)), [1.0, 0.5, 0.75, 2.0 ** -30, 3.0]]):
        s, inv = ol.h_scale(n), ol.h_scale(n, True)
        assert 0.5 <= n * s < 1.0
        assert s * inv == 1.0 and np.log2(s) == np.round(np.log2(s))
    assert ol.h_scale(0.0) == 1.0 and ol.h_scale(0.0, True) == 1.0


# ------------------------------------------------------------ operators

SHAPES = [(9, 9, 9), (9, 17, 5), (17, 9, 13), (5, 5, 5)]


def eighths(rng, shape, lim=16):
    return (rng.integers(-lim, lim + 1, size=shape) / 8.0)


def interior_box(shape):
    return [1, shape[0] - 1, 1, shape[1] - 1, 1, shape[2] - 1]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("colour", [0, 1])
def test_h_gs_half_exact_inputs_equal_fp64_operator(shape, colour):
    rng = np.random.default_rng(sum(shape) + colour)
    u, f = eighths(rng, shape), eighths(rng, shape)
    box = interior_box(shape)
    u64 = u.copy()
    L = ol.layout(shape)
    ol.lib().og_gs_half(ol.CONSISTENT, L, ol.ptr(u64), ol.ptr(f), ol.box_arr(box), 0.75, 0.5, colour)
    u16 = u.astype(np.float16)
    assert np.array_equal(u16.astype(np.float64), u)  # inputs exact in binary16
    ol.h_gs_half(u16, f.astype(np.float16), box, 0.75, 0.5, colour)
    assert np.array_equal(u16.astype(np.float64), u64)  # results exact too: equal, not close
    assert not np.array_equal(u64, u)


@pytest.mark.parametrize("shape", SHAPES)
def test_h_residual_exact_inputs_equal_fp64_operator(shape):
    rng = np.random.default_rng(sum(shape))
    u, f = eighths(rng, shape), eighths(rng, shape)
    box = interior_box(shape)
    r64 = np.zeros(shape)
    ol.lib().og_residual(ol.CONSISTENT, ol.layout(shape), ol.ptr(u), ol.ptr(f), ol.ptr(r64), ol.box_arr(box), 0.5)
    r16 = np.zeros(shape, dtype=np.float16)
    m = ol.h_residual(u.astype(np.float16), f.astype(np.float16), r16, box, 0.5)
    assert np.array_equal(r16.astype(np.float64), r64)
    assert m == np.abs(r64).max() > 0


@pytest.mark.parametrize("fine", [(9, 9, 9), (17, 9, 13), (5, 9, 17)])
@pytest.mark.parametrize("src", [np.float16, np.float32])
@pytest.mark.parametrize("norm", [0.0, 3.0, 0.01])
def test_h_restrict_fw_exact_inputs_equal_fp64_operator(fine, src, norm):
    rng = np.random.default_rng(sum(fine))
    a = eighths(rng, fine)
    coarse = tuple((n - 1) // 2 + 1 for n in fine)
    c64 = np.zeros(coarse)
    ol.lib().og_restrict_fw(ol.layout(fine), ol.ptr(a), ol.layout(coarse), ol.ptr(c64))
    c16 = np.zeros(coarse, dtype=np.float16)
    ol.h_restrict_fw(a.astype(src), c16, norm)
    sigma = {0.0: 1.0, 3.0: 0.25, 0.01: 64.0}[norm]  # 2^-e, norm = m 2^e, m in [0.5, 1)
    assert np.array_equal(c16.astype(np.float64), c64 * sigma)


@pytest.mark.parametrize("coarse", [(5, 5, 5), (9, 5, 7), (3, 9, 5)])
@pytest.mark.parametrize("dst", [np.float16, np.float32])
@pytest.mark.parametrize("norm", [0.0, 3.0, 0.3])
def test_h_prolong_add_exact_inputs_equal_fp64_operator(coarse, dst, norm):
    rng = np.random.default_rng(sum(coarse))
    c = eighths(rng, coarse)
    fine = tuple(2 * (n - 1) + 1 for n in coarse)
    p64 = np.zeros(fine)
    ol.lib().og_prolong_trilinear(ol.layout(coarse), ol.ptr(c), ol.layout(fine), ol.ptr(p64))
    u = eighths(rng, fine)
    out = u.astype(dst)
    ol.h_prolong_add(c.astype(np.float16), out, norm)
    inv = {0.0: 1.0, 3.0: 4.0, 0.3: 0.5}[norm]  # sums stay exact in binary16
    assert np.array_equal(out.astype(np.float64), u + p64 * inv)


def half_ulp(x):
    """Half a binary16 ulp of |x| (normal range), the rounding bound."""
    e = np.floor(np.log2(np.maximum(np.abs(x), 2.0 ** -14)))
    return np.exp2(e - 11)


@pytest.mark.parametrize("shape", SHAPES)
def test_h_operators_random_inputs_within_rounding_bound(shape):
    rng = np.random.default_rng(7 + sum(shape))
    u16 = rng.standard_normal(shape).astype(np.float16)
    f16 = rng.standard_normal(shape).astype(np.float16)
    u, f = u16.astype(np.float64), f16.astype(np.float64)
    box = interior_box(shape)
    h, omega = 1.0 / 16, 8.0 / 9.0
    L = ol.layout(shape)
    # smoother
    u64 = u.copy()
    ol.lib().og_gs_half(ol.CONSISTENT, L, ol.ptr(u64), ol.ptr(f), ol.box_arr(box), omega, h, 0)
    g16 = u16.copy()
    ol.h_gs_half(g16, f16, box, omega, h, 0)
    err = np.abs(g16.astype(np.float64) - u64)
    assert np.all(err <= half_ulp(u64) + 1e-6 * (np.abs(u64) + 1))
    # residual (float rounding of the h^-2-scaled sum dominates the second term)
    r64 = np.zeros(shape)
    ol.lib().og_residual(ol.CONSISTENT, L, ol.ptr(u), ol.ptr(f), ol.ptr(r64), ol.box_arr(box), h)
    r16 = np.zeros(shape, dtype=np.float16)
    ol.h_residual(u16, f16, r16, box, h)
    err = np.abs(r16.astype(np.float64) - r64)
    assert np.all(err <= half_ulp(r64) + 1e-6 * (np.abs(r64) + 16 * 256))
    assert np.sum(err > 0.5 * half_ulp(r64)) > 0  # rounding happened somewhere
