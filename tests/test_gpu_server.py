"""Serving loop (mgp_server_*) on the GPU: every request's result is the same
bits when it runs alone and when it runs among others, each reaches the
tolerance, and admission respects the stream-slot limit."""
import zlib

import numpy as np
import pytest

import mgp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def requests(seed, count, kmax=6):
    """Shapes 2^k+1 per axis (k in 3..kmax), face types by seed (never all
    Neumann), four Fourier modes."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        dims = tuple(int(2 ** k + 1) for k in rng.integers(3, kmax + 1, size=3))
        bc = "".join("N" if v else "D" for v in rng.integers(0, 2, size=6))
        if "D" not in bc:
            continue
        modes = [(int(a), int(b), int(c), float(rng.standard_normal())) for a, b, c in rng.integers(1, 9, size=(4, 3))]
        out.append((dims, bc, modes))
    return out


def test_results_alone_equal_results_in_the_mix():
    reqs = requests(zlib.crc32(b"serve"), 10)
    say("case server alone")
    alone = []
    s = mgp.Server(max_active=1, tolerance=1e-8)
    for dims, bc, modes in reqs:
        res, u = s.wait(s.submit(dims, bc, modes))
        assert res.status == 0
        assert res.rel_res <= 1e-8, (dims, bc, res.rel_res, res.iterations)
        alone.append((res.iterations, res.rel_res, u))
    s.close()
    say("case server mix")
    s = mgp.Server(max_active=4, tolerance=1e-8)
    ids = [s.submit(dims, bc, modes) for dims, bc, modes in reqs]
    for k, rid in enumerate(ids):
        res, u = s.wait(rid)
        assert res.status == 0
        assert (res.iterations, res.rel_res) == alone[k][:2], (reqs[k][:2], res.iterations, alone[k][:2])
        np.testing.assert_array_equal(u, alone[k][2])
    completed, peak, resident = s.stats()
    assert completed == len(reqs)
    assert 1 <= peak <= 4
    say("peak active", peak, "resident bytes", resident)
    s.close()


def test_tight_budget_still_serves_every_request():
    """A budget that holds about one hierarchy forces eviction of idle
    hierarchies and admission one at a time; results stay the same."""
    reqs = requests(zlib.crc32(b"budget"), 5, kmax=5)
    s = mgp.Server(max_active=4, hbm_budget=16 << 20, tolerance=1e-8)
    ids = [s.submit(dims, bc, modes) for dims, bc, modes in reqs]
    got = [s.wait(i) for i in ids]
    s.close()
    ref = mgp.Server(max_active=1, tolerance=1e-8)
    for (dims, bc, modes), (res, u) in zip(reqs, got):
        r2, u2 = ref.wait(ref.submit(dims, bc, modes))
        assert res.status == 0 and (res.iterations, res.rel_res) == (r2.iterations, r2.rel_res)
        np.testing.assert_array_equal(u, u2)
    ref.close()


def test_rejects_all_neumann_and_bad_shapes():
    s = mgp.Server(max_active=1)
    with pytest.raises(mgp.MgpError):
        s.submit((17, 17, 17), "NNNNNN", [(1, 1, 1, 1.0)])
    with pytest.raises(mgp.MgpError):
        s.submit((2, 17, 17), "DDDDDD", [(1, 1, 1, 1.0)])
    s.close()
