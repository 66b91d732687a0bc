#!/usr/bin/env python3
"""Record golden vectors from the reference itself.

Runs oracle/_ref/ref_harness (the reference library compiled unmodified from
its own sources by `make -C oracle ref`) on small cases and stores inputs and
outputs as .npz fixtures in this directory.  Each fixture holds a JSON
"params" entry plus numpy arrays (no pickles).

Run in the development container only (the reference tree is not present on
the GPU machines; the fixtures travel instead):

    make -C oracle ref && python3 tests/golden/make_goldens.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

LOC = "NSEWTB"


def halo_of(bc):
    return [1 if c == "N" else 0 for c in bc]


def padded(n, halo):
    hN, hS, hE, hW, hT, hB = halo
    return (n[0] + hE + hW, n[1] + hN + hS, n[2] + hT + hB)


def face_shape(loc, n):
    fs = list(n)
    if loc in (2, 3):
        fs[0] = 1
    if loc in (0, 1):
        fs[1] = 1
    if loc in (4, 5):
        fs[2] = 1
    return tuple(fs)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(params, inputs=None, threads=4):
    """Run the harness; returns the out-dir contents as raw bytes per file."""
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        with open(os.path.join(td, "p.txt"), "w") as fh:
            for k, v in params.items():
                fh.write("%s %s\n" % (k, v))
        for name, arr in (inputs or {}).items():
            np.ascontiguousarray(arr).tofile(os.path.join(ind, name + ".bin"))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        subprocess.run([HARNESS, os.path.join(td, "p.txt"), ind, outd], check=True, env=env,
                       stdout=subprocess.DEVNULL)
        return {f[:-4]: open(os.path.join(outd, f), "rb").read() for f in os.listdir(outd)}


def solve_case(name, dims, levels, bc, problem, restriction, relaxation, cycle, nsmooth, omega,
               ncycles, lengthx=1.0, origin=(0.0, 0.0, 0.0), dtype="f64", threads=4, full=True,
               init_zero=0):
    params = dict(mode="solve", dtype=dtype, nx=dims[0], ny=dims[1], nz=dims[2], levels=levels,
                  lengthx=repr(float(lengthx)), ox=repr(float(origin[0])), oy=repr(float(origin[1])),
                  oz=repr(float(origin[2])), bc=bc, problem=problem, restriction=restriction,
                  relaxation=relaxation, cycle=cycle, nsmooth=nsmooth, omega=repr(float(omega)),
                  ncycles=ncycles, init_zero=init_zero)
    out = run(params, threads=threads)
    dt = np.float64 if dtype == "f64" else np.float32
    shp = padded(dims, halo_of(bc))
    u = np.frombuffer(out["u"], dtype=dt).reshape(shp)
    r = np.frombuffer(out["r"], dtype=dt).reshape(shp)
    f = np.frombuffer(out["f"], dtype=dt).reshape(shp)
    rec = dict(params=json.dumps(params), relres=np.frombuffer(out["relres"], np.float64),
               maxerr=np.frombuffer(out["maxerr"], np.float64), u_sha=sha(u), r_sha=sha(r),
               f_sha=sha(f), u_sub=u[::4, ::4, ::4].copy())
    if full:
        rec["u"] = u
        rec["r"] = r
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
    print("%-28s relres[-1]=%.6e maxerr[-1]=%.3e" % (name, rec["relres"][-1], rec["maxerr"][-1]))


def stop_case(name, dims, levels, bc, problem, restriction, relaxation, cycle, nsmooth, omega, maxiter,
              miniter, tol):
    """One solve() with the reference's own stop rule (libpoisson.h:340-353)."""
    params = dict(mode="solve", dtype="f64", nx=dims[0], ny=dims[1], nz=dims[2], levels=levels,
                  lengthx="1.0", ox="0.0", oy="0.0", oz="0.0", bc=bc, problem=problem,
                  restriction=restriction, relaxation=relaxation, cycle=cycle, nsmooth=nsmooth,
                  omega=repr(float(omega)), full_solve=1, maxiter=maxiter, miniter=miniter,
                  tol=repr(float(tol)))
    out = run(params)
    stop = np.frombuffer(out["stop"], np.float64)
    u = np.frombuffer(out["u"], np.float64)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), params=json.dumps(params),
                        iterations=stop[0], rel_res=stop[1], u_sha=sha(u))
    print("%-28s iterations=%d rel_res=%.6e" % (name, stop[0], stop[1]))


def rand(rng, shape):
    return rng.uniform(-1.0, 1.0, size=shape)


def op_level_case(name, op, dims, bc, seed, extra=None, threads=4):
    """residual / relax_gs / relax_jacobi / gs_kernel on a random one-level domain."""
    rng = np.random.default_rng(seed)
    halo = halo_of(bc)
    shp = padded(dims, halo)
    inputs = {k: rand(rng, shp) for k in ("u", "uprev", "f", "r")}
    for q in range(6):
        inputs["g" + LOC[q]] = rand(rng, face_shape(q, dims))
    params = dict(mode="op", op=op, dtype="f64", nx=dims[0], ny=dims[1], nz=dims[2], lengthx="1.0",
                  ox="0.0", oy="0.0", oz="0.0", bc=bc)
    params.update(extra or {})
    out = run(params, inputs, threads=threads)
    rec = {"params": json.dumps(params)}
    rec.update({"in_" + k: v for k, v in inputs.items()})
    for k, v in out.items():
        rec["out_" + k] = np.frombuffer(v, np.float64).reshape(shp)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
    print("%-28s %s" % (name, op))


def op_transfer_case(name, op, n_in, n_out, halo_in, halo_out, seed):
    rng = np.random.default_rng(seed)
    a = rand(rng, padded(n_in, halo_in))
    b = rand(rng, padded(n_out, halo_out))
    params = dict(mode="op", op=op, dtype="f64", nx=n_in[0], ny=n_in[1], nz=n_in[2], lengthx="1.0",
                  ox="0.0", oy="0.0", oz="0.0")
    for pre, n, h in (("in", n_in, halo_in), ("out", n_out, halo_out)):
        params.update({pre + "_nx": n[0], pre + "_ny": n[1], pre + "_nz": n[2]})
        params.update({pre + "_h" + LOC[q]: h[q] for q in range(6)})
    out = run(params, {"in": a, "out": b})
    rec = dict(params=json.dumps(params), in_a=a, in_b=b,
               out_b=np.frombuffer(out["out"], np.float64).reshape(b.shape))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
    print("%-28s %s" % (name, op))


def op_bc_restrict_case(name, dims, bc, restriction, seed):
    rng = np.random.default_rng(seed)
    halo = halo_of(bc)
    nc = tuple(n // 2 + 1 for n in dims)
    inputs = {"u": rand(rng, padded(dims, halo))}
    for q in range(6):
        inputs["g" + LOC[q]] = rand(rng, face_shape(q, dims))
        inputs["cg" + LOC[q]] = rand(rng, face_shape(q, nc))
    lengthx = 1.0
    params = dict(mode="op", op="bc_restrict", dtype="f64", nx=dims[0], ny=dims[1], nz=dims[2],
                  lengthx=repr(lengthx), ox="0.0", oy="0.0", oz="0.0", bc=bc, restriction=restriction)
    out = run(params, inputs)
    rec = {"params": json.dumps(params)}
    rec.update({"in_" + k: v for k, v in inputs.items()})
    for q in range(6):
        rec["out_cg" + LOC[q]] = np.frombuffer(out["cg" + LOC[q]], np.float64).reshape(face_shape(q, nc))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
    print("%-28s bc_restrict %s" % (name, restriction))


def vtk_case(name, nx, ny, nz, origin, h):
    """The reference's VTK writers (include/array.h:148-269) on a fixed field,
    the files kept byte for byte."""
    params = dict(mode="vtk", nx=nx, ny=ny, nz=nz, ox=repr(float(origin[0])), oy=repr(float(origin[1])),
                  oz=repr(float(origin[2])), h=repr(float(h)))
    out = run(params)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), params=json.dumps(params),
                        vtk=np.frombuffer(out["u"], np.uint8), vtk_halo=np.frombuffer(out["u_halo"], np.uint8))
    print("%-28s %d + %d bytes" % (name, len(out["u"]), len(out["u_halo"])))


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    # G1: 65^3, 4 levels, V(2,2) red-black GS w=1, injection, all Dirichlet,
    # u = sin(2 pi x) sin(2 pi y) sin(2 pi z) on [0,1]^3, 10 cycles
    solve_case("g1_c1_vcycle_gs_inj_65", (65, 65, 65), 4, "DDDDDD", 0, "inj", "gs", "Vcycle", 2, 1.0, 10,
               full=False)
    # G2: 33^3, F-cycle, GS (default omega), full weighting, all Dirichlet, problem 3
    solve_case("g2_fcycle_gs_fw_p3_33", (33, 33, 33), 3, "DDDDDD", 3, "fw", "gs", "Fcycle", 2, -1.0, 6)
    # G3: 33x33x17, F-cycle, GS, FW, requested {N,N,N,N,D,D} (north/south forced
    # Dirichlet, Neumann east/west with the reference's coarse Neumann data), problem 2
    solve_case("g3_fcycle_gs_fw_neumann_p2", (33, 33, 17), 3, "NNNNDD", 2, "fw", "gs", "Fcycle", 2, 1.0, 6)
    # G4: 17^3 Jacobi V-cycle, injection, east/west Neumann, reference default
    # origin and length (settings.h:23-29), problem 1
    solve_case("g4_vcycle_jacobi_inj_ew_neumann", (17, 17, 17), 3, "DDNNDD", 1, "inj", "jacobi", "Vcycle", 3,
               -1.0, 6, lengthx=3.1415, origin=(-0.4, 0.21, 1.23))
    # G5: 17^3 GS, top Neumann / bottom Dirichlet: even z stride, the
    # linear-index colouring degenerates to z-planes (a data race in the
    # reference; recorded with one thread = sequential order)
    solve_case("g5_degenerate_colouring_1thread", (17, 17, 17), 3, "DDDDND", 3, "inj", "gs", "Vcycle", 2, 1.0,
               4, threads=1)
    # G6: F-cycle, Jacobi, FW, all faces requested Neumann (drivers/neumann.cpp)
    solve_case("g6_fcycle_jacobi_fw_allN", (17, 17, 17), 3, "NNNNNN", 1, "fw", "jacobi", "Fcycle", 2, -1.0, 5,
               lengthx=1.0, origin=(0.1, 0.2, 0.3))
    # G7: single precision (PoissonSolver<float,...>)
    solve_case("g7_float_vcycle_gs_fw_33", (33, 33, 33), 3, "DDDDDD", 0, "fw", "gs", "Vcycle", 2, 1.0, 6,
               dtype="f32")
    # G8: "relaxation" solver: no work per iteration (libpoisson.h:304-308)
    solve_case("g8_relaxation_noop", (9, 9, 9), 2, "DDDDDD", 0, "inj", "gs", "relaxation", 2, 1.0, 3)
    # G9: 2 levels, V-cycle with default GS omega 8/9, non-cubic grid
    solve_case("g9_vcycle_gs_2level_noncubic", (17, 9, 33), 2, "DDNNNN", 2, "fw", "gs", "Vcycle", 3, -1.0, 5)
    # G10: init_zero() path (all data zero) -> rel_res is 0/0
    # (not recorded: NaN semantics are host-library dependent)

    # the reference's stop rule: divergence break (G1 grows 3x per cycle),
    # stagnation break |norm - prev| < tol, and running to maxiter
    stop_case("stop_divergence_c1", (65, 65, 65), 4, "DDDDDD", 0, "inj", "gs", "Vcycle", 2, 1.0, 10, 1, 1e-6)
    stop_case("stop_stagnation_neumann", (33, 33, 17), 3, "NNNNDD", 2, "fw", "gs", "Fcycle", 2, 1.0, 40, 1, 1e-4)
    stop_case("stop_maxiter_fw", (17, 17, 17), 3, "DDDDDD", 3, "fw", "gs", "Vcycle", 2, 1.0, 4, 1, 1e-14)
    stop_case("stop_miniter3", (17, 17, 17), 3, "DDDDDD", 3, "fw", "jacobi", "Fcycle", 2, -1.0, 30, 3, 1e-3)

    # single operators on random inputs
    op_level_case("op_residual_dirichlet_9", "residual", (9, 9, 9), "DDDDDD", 1)
    op_level_case("op_residual_mixed_9x11x7", "residual", (9, 11, 7), "NNNNND", 2)
    op_level_case("op_relax_gs_mixed_9", "relax_gs", (9, 9, 9), "NNNNDD", 3, dict(omega="0.9", nsmooth=2))
    op_level_case("op_relax_gs_topN_9", "relax_gs", (9, 9, 9), "DDDDND", 4, dict(omega="1.1", nsmooth=1),
                  threads=1)
    op_level_case("op_relax_jacobi_mixed_9", "relax_jacobi", (9, 9, 9), "NNNNNN", 5,
                  dict(omega="0.857142857142857", nsmooth=3))
    op_level_case("op_gs_kernel_11x9x13", "gs_kernel", (11, 9, 13), "DDNNDD", 6, dict(omega="1.0", halfsweeps=3))
    op_transfer_case("op_restrict_inj_17_9", "restrict_inj", (17, 17, 17), (9, 9, 9), [0, 0, 1, 1, 0, 0],
                     [0, 0, 1, 1, 0, 0], 7)
    op_transfer_case("op_restrict_fw_17x9x13", "restrict_fw", (17, 9, 13), (9, 5, 7), [1, 1, 1, 1, 1, 0],
                     [1, 1, 1, 1, 1, 0], 8)
    op_transfer_case("op_restrict_fw_face_1x17x9", "restrict_fw", (1, 17, 9), (1, 9, 5), [0] * 6, [0] * 6, 9)
    op_transfer_case("op_prolong_9x5x7", "prolong", (9, 5, 7), (17, 9, 13), [0, 0, 1, 1, 0, 1],
                     [0, 0, 1, 1, 0, 1], 10)
    op_bc_restrict_case("op_bc_restrict_fw_17", (17, 17, 17), "NNNNNN", "fw", 11)
    op_bc_restrict_case("op_bc_restrict_inj_17x9x9", (17, 9, 9), "DNNNND", "inj", 12)
    vtk_case("fmt_vtk_3x4x5", 3, 4, 5, (0.5, -1.0, 2.0), 0.25)


if __name__ == "__main__":
    main()
