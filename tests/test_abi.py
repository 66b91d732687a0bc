"""The C-ABI library and the drop-in headers, checked on the CPU (no GPU calls)."""
import os
import subprocess

import pytest

import mgp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "multigrid-poisson_amd")
REFERENCE = "/root/reference"


def test_library_loads_and_exports_every_declared_symbol():
    L = mgp.lib()
    names = mgp.declared_symbols()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    bound = set(mgp._SIGS) | {"mgp_error_string", "mgp_version", "mgp_last_error", "mgp_solver_init_fn"}
    assert set(mgp.declared_symbols()) <= bound


def test_version_and_error_strings():
    L = mgp.lib()
    assert b"gfx950" in L.mgp_version()
    assert L.mgp_error_string(-1) == b"invalid argument"


def test_invalid_arguments_are_rejected_without_touching_a_gpu():
    L = mgp.lib()
    a = mgp.Array()  # data == NULL
    assert L.mgp_fill_f64(a, 0.0, None) == -1
    assert L.mgp_restrict_fw_f64(a, a, None) == -1
    assert L.mgp_solver_create(None, None) == -1
    cfg = mgp.SolverConfig()
    cfg.bc[0] = 7
    h = mgp.vp()
    assert L.mgp_solver_create(cfg, h) == -1


def _gxx(src, out, extra=()):
    cmd = ["g++", "-std=c++17", "-O0", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(PKG, "poisson"), "-I" + os.path.join(PKG, "csrc")] + list(extra) + [src, "-o", out,
           "-L" + os.path.join(PKG, "lib"), "-lmgp"]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_headers_compile_with_plain_gxx(tmp_path):
    """The Poisson:: headers are plain C++ over the C ABI: g++ builds a program
    that uses every policy combination without any HIP compiler."""
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include "libpoisson.h"
namespace Poisson {
double_t ufun(double_t x, double_t y, double_t z) { return x * y * z; }
double_t ffun(double_t, double_t, double_t) { return 0.0; }
double_t dudxfun(double_t, double_t y, double_t z) { return y * z; }
double_t dudyfun(double_t x, double_t, double_t z) { return x * z; }
double_t dudzfun(double_t x, double_t y, double_t) { return x * y; }
}
using namespace Poisson;
template <class T, template <class> class R, template <class> class S> void use(Settings& s, BoundaryCondition& bc) {
    PoissonSolver<T, R, TrilinearInterpolation, S> solver(s, bc);
    solver.init(); solver.init_zero(); solver.to_device(); solver.solve("Fcycle", 2, 1.0); solver.to_host();
    (void)solver.get_solution(0); (void)solver.relative_residual(); (void)solver.get_top_bc(0);
}
int main() {
    Settings s; BoundaryCondition bc{NEUMANN, NEUMANN, NEUMANN, NEUMANN, DIRICHLET, DIRICHLET};
    use<double, Injection, GaussSeidel>(s, bc); use<double, FullWeighting, Jacobi>(s, bc);
    use<float, Injection, Jacobi>(s, bc); use<float, FullWeighting, GaussSeidel>(s, bc);
    return 0;
}
''')
    r = _gxx(str(src), str(tmp_path / "t"))
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree not present")
@pytest.mark.parametrize("driver", ["dirichlet", "neumann", "mixed", "gaussseidel", "vcycle", "fcycle"])
def test_reference_drivers_compile_against_the_drop_in_headers(driver, tmp_path):
    """The reference's own example programs build unchanged against these
    headers (their problem definitions come from the reference's
    src/problem_definition.cpp, as with the reference's own build)."""
    src = os.path.join(REFERENCE, "drivers", driver + ".cpp")
    prob = os.path.join(REFERENCE, "src", "problem_definition.cpp")
    obj = tmp_path / "p.o"
    r1 = subprocess.run(["g++", "-std=c++17", "-c", "-DPROBLEM2", "-I" + os.path.join(PKG, "poisson"),
                         "-I" + os.path.join(ROOT, "include"), prob, "-o", str(obj)], capture_output=True, text=True)
    assert r1.returncode == 0, r1.stderr[-3000:]
    r = _gxx(src, str(tmp_path / driver), ["-fopenmp", "-DPROBLEM2", str(obj)])
    assert r.returncode == 0, r.stderr[-4000:]


def test_level_settings_follow_the_coarsening_rule(tmp_path):
    """DomainSettings / Grid (the reference's src/domainsettings.cpp:12-29
    rule): level l of n points per axis has (n >> l) + 1 points and spacing
    h 2^l; an axis that is not 2^l m + 1 throws invalid_argument naming the
    coarse size, as the reference's message does."""
    src = tmp_path / "g.cpp"
    src.write_text(r'''
#include <cstdio>
#include "grid.h"
using namespace Poisson;
int main() {
    Settings s; s.dims[0] = 1025; s.dims[1] = 513; s.dims[2] = 33; s.h = 1.0 / 1024;
    Grid g(s, 6, false);
    for (int l = 0; l < 6; l++)
        std::printf("%ld %ld %ld %.17g\n", g.domainsettings[l].dims[0], g.domainsettings[l].dims[1],
                    g.domainsettings[l].dims[2], g.domainsettings[l].h);
    s.dims[2] = 35;
    try { DomainSettings d(s, 2); std::printf("no throw\n"); }
    catch (const std::invalid_argument& e) { std::printf("%s\n", e.what()); }
    s.dims[2] = 34;
    try { DomainSettings d(s, 1); std::printf("no throw\n"); }
    catch (const std::invalid_argument& e) { std::printf("%s\n", e.what()); }
    s.dims[2] = 37;
    try { DomainSettings d(s, 2); std::printf("no throw\n"); }
    catch (const std::invalid_argument& e) { std::printf("%s\n", e.what()); }
    return 0;
}
''')
    exe = tmp_path / "g"
    r = subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "poisson"),
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines()
    want = []
    for l in range(6):
        n = [(m >> l) + 1 if l else m for m in (1025, 513, 33)]
        want.append("%d %d %d %.17g" % (n[0], n[1], n[2], 2.0 ** l / 1024))
    assert out[:6] == want
    assert out[6] == "Invalid grid of size: (257,129,9)"
    assert out[7] == "Invalid grid of size: (513,257,18)"
    assert out[8] == "no throw"
