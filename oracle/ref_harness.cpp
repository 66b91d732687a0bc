// Test infrastructure only: a driver that runs the reference library
// (eljens/multigrid-poisson, header templates + src/*.cpp, compiled unmodified
// from the reference tree) on the host so that its outputs can be recorded as
// golden vectors.  Nothing in the product links or calls this file.
//
// Two pieces make the reference runnable on a GPU-less host:
//  * omp_get_num_devices() is interposed to return 1, because
//    include/devicearray.h:117-119 throws when the OpenMP runtime reports no
//    device.  With no offload plugin, libgomp treats device 0 as the host
//    (omp_target_alloc = malloc, omp_target_memcpy = memcpy, target regions
//    fall back to the host), so every kernel runs as plain host code.
//  * the problem functions declared in include/problem_definition.h are
//    defined here (the reference selects them at compile time with
//    -DPROBLEMn in src/problem_definition.cpp); a run-time id picks one.
//
// Usage:  ref_harness <params-file> <in-dir> <out-dir>
//         (mode "time": the reference's own solve timed; in-dir and out-dir unused)
// The params file is "key value" lines; arrays are raw little-endian buffers in
// the reference's own halo-padded layout (Array<T>::at, include/array.h:84-94).
#include "libpoisson.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

extern "C" int omp_get_num_devices(void) { return 1; }

namespace {
int g_problem = 0;
}

namespace Poisson {
// Problem 0: u = sin(2 pi x) sin(2 pi y) sin(2 pi z), f = -12 pi^2 u.
// Problems 1-3 restate src/problem_definition.cpp:1-76 (PROBLEM1..3).
double_t ufun(double_t x, double_t y, double_t z) {
    switch (g_problem) {
    case 0: { const double p = 2.0 * M_PI; return sin(p * x) * sin(p * y) * sin(p * z); }
    case 1: return sin(3.0 * x) * sin(5.3 * y) * sin(-7.8 * z);
    case 2: return x * x * x * y * y * z;
    default: return cos(x * z * z) * sin(y * y * y);
    }
}
double_t ffun(double_t x, double_t y, double_t z) {
    switch (g_problem) {
    case 0: { const double p = 2.0 * M_PI; return -3.0 * p * p * sin(p * x) * sin(p * y) * sin(p * z); }
    case 1: return -(3.0 * 3.0 + 5.3 * 5.3 + -7.8 * -7.8) * sin(3.0 * x) * sin(5.3 * y) * sin(-7.8 * z);
    case 2: return 2.0 * x * x * x * z + 6.0 * x * y * y * z;
    default:
        return ((-4.0 * x * x * z * z - 9.0 * y * y * y * y - z * z * z * z) * sin(y * y * y) +
                6.0 * y * cos(y * y * y)) * cos(x * z * z) - 2.0 * x * sin(x * z * z) * sin(y * y * y);
    }
}
double_t dudxfun(double_t x, double_t y, double_t z) {
    switch (g_problem) {
    case 0: { const double p = 2.0 * M_PI; return p * cos(p * x) * sin(p * y) * sin(p * z); }
    case 1: return 3.0 * cos(3.0 * x) * sin(5.3 * y) * sin(-7.8 * z);
    case 2: return 3.0 * x * x * y * y * z;
    default: return -z * z * sin(x * z * z) * sin(y * y * y);
    }
}
double_t dudyfun(double_t x, double_t y, double_t z) {
    switch (g_problem) {
    case 0: { const double p = 2.0 * M_PI; return p * sin(p * x) * cos(p * y) * sin(p * z); }
    case 1: return 5.3 * sin(3.0 * x) * cos(5.3 * y) * sin(-7.8 * z);
    case 2: return 2.0 * x * x * x * y * z;
    default: return 3.0 * cos(x * z * z) * y * y * cos(y * y * y);
    }
}
double_t dudzfun(double_t x, double_t y, double_t z) {
    switch (g_problem) {
    case 0: { const double p = 2.0 * M_PI; return p * sin(p * x) * sin(p * y) * cos(p * z); }
    case 1: return -7.8 * sin(3.0 * x) * sin(5.3 * y) * cos(-7.8 * z);
    case 2: return x * x * x * y * y;
    default: return -2 * x * z * sin(x * z * z) * sin(y * y * y);
    }
}
}  // namespace Poisson

using namespace Poisson;

namespace {

struct Params {
    std::map<std::string, std::string> kv;
    std::string s(const std::string& k, const std::string& d = "") const {
        auto it = kv.find(k);
        return it == kv.end() ? d : it->second;
    }
    long i(const std::string& k, long d = 0) const {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::stol(it->second);
    }
    double f(const std::string& k, double d = 0) const {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::stod(it->second);
    }
};

Params read_params(const char* path) {
    Params p;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string k, v;
        if (ss >> k) {
            std::getline(ss, v);
            size_t b = v.find_first_not_of(" \t");
            p.kv[k] = b == std::string::npos ? "" : v.substr(b);
        }
    }
    return p;
}

template <class T>
void load(const std::string& path, T* dst, size_t n) {
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) { fprintf(stderr, "missing input %s\n", path.c_str()); exit(2); }
    if (fread(dst, sizeof(T), n, fp) != n) { fprintf(stderr, "short input %s\n", path.c_str()); exit(2); }
    fclose(fp);
}

template <class T>
void save(const std::string& path, const T* src, size_t n) {
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(2); }
    fwrite(src, sizeof(T), n, fp);
    fclose(fp);
}

BoundaryCondition parse_bc(const std::string& s) {
    // order: north south east west top bottom, 'D' or 'N'
    auto t = [&](int i) { return s.at(i) == 'N' ? NEUMANN : DIRICHLET; };
    BoundaryCondition bc;
    bc.north = t(0); bc.south = t(1); bc.east = t(2);
    bc.west = t(3); bc.top = t(4); bc.bottom = t(5);
    return bc;
}

Settings make_settings(const Params& p) {
    Settings s;
    s.dims[0] = p.i("nx"); s.dims[1] = p.i("ny"); s.dims[2] = p.i("nz");
    s.levels = p.i("levels", 1);
    s.lengthx = p.f("lengthx", 1.0);
    s.h = s.lengthx / ((double_t)s.dims[0] - 1.0);  // src/parser.cpp:48
    s.origin[0] = p.f("ox"); s.origin[1] = p.f("oy"); s.origin[2] = p.f("oz");
    s.maxiter = 1;
    s.miniter = 1;
    s.tolerance = 0.0;
    return s;
}

// Full solve through PoissonSolver.  solve() is called once per cycle with
// maxiter = 1 so that rel_res can be read after every cycle at full
// precision; with maxiter = 1 the stop rule never fires and the extra initial
// residual each call recomputes r from an unchanged u (no state change).
template <class T, template <class> class R, template <class> class S>
int run_solve(const Params& p, const std::string& outdir) {
    Settings settings = make_settings(p);
    if (p.i("full_solve", 0)) {
        settings.maxiter = p.i("maxiter");
        settings.miniter = p.i("miniter");
        settings.tolerance = p.f("tol");
    }
    BoundaryCondition bc = parse_bc(p.s("bc", "DDDDDD"));
    PoissonSolver<T, R, TrilinearInterpolation, S> solver(settings, bc);
    if (p.i("init_zero", 0)) {
        solver.init_zero();
    } else {
        solver.init();
    }
    solver.to_device();
    const long ncycles = p.i("ncycles");
    const long nsmooth = p.i("nsmooth", 10);
    const double omega = p.f("omega", -1.0);
    const std::string cycle = p.s("cycle", "Vcycle");
    std::vector<double> relres, maxerr;
    auto err_now = [&]() {
        const DeviceArray<T>& u = solver.get_solution(0);
        double e = 0.0;
        for (uint_t i = 0; i < u.shape[0]; i++)
            for (uint_t j = 0; j < u.shape[1]; j++)
                for (uint_t k = 0; k < u.shape[2]; k++) {
                    double ut = ufun(settings.origin[0] + ((double)i) * settings.h,
                                     settings.origin[1] + ((double)j) * settings.h,
                                     settings.origin[2] + ((double)k) * settings.h);
                    double d = std::abs((double)u.at[u.idx(i, j, k)] - ut);
                    e = std::max(e, d);
                }
        return e;
    };
    if (p.i("full_solve", 0)) {
        // one solve() call with the caller's maxiter/miniter/tolerance, which
        // exercises the reference's own stop rule (libpoisson.h:340-353)
        solver.solve(cycle, nsmooth, (T)omega);
        solver.to_host();
        double rec[2] = {(double)solver.solve_iterations(), (double)solver.relative_residual()};
        save(outdir + "/stop.bin", rec, 2);
        const DeviceArray<T>& us = solver.get_solution(0);
        save(outdir + "/u.bin", us.at, us.size);
        return 0;
    }
    for (long c = 0; c < ncycles; c++) {
        solver.solve(cycle, nsmooth, (T)omega);
        solver.to_host();
        relres.push_back((double)solver.relative_residual());
        maxerr.push_back(err_now());
        if (p.i("dump_every", 0)) {
            const DeviceArray<T>& uu = solver.get_solution(0);
            save(outdir + "/u_" + std::to_string(c) + ".bin", uu.at, uu.size);
        }
    }
    save(outdir + "/relres.bin", relres.data(), relres.size());
    save(outdir + "/maxerr.bin", maxerr.data(), maxerr.size());
    const DeviceArray<T>& uf = solver.get_solution(0);
    save(outdir + "/u.bin", uf.at, uf.size);
    const DeviceArray<T>& rf = solver.get_residual(0);
    save(outdir + "/r.bin", rf.at, rf.size);
    DeviceArray<T>& ff = solver.get_rhs(0);
    save(outdir + "/f.bin", ff.at, ff.size);
    return 0;
}

template <class T>
struct OneLevel {
    Settings settings;
    BoundaryCondition bc;
    Domain<T>* dom;
    Domain<T>** levels[1];
    OneLevel(const Params& p, bool dup) {
        settings = make_settings(p);
        bc = parse_bc(p.s("bc", "DDDDDD"));
        int_t dev[3] = {0, 0, 0};
        int_t shape[3] = {1, 1, 1};
        dom = new Domain<T>(settings, bc, dup, dev, shape);
        levels[0] = &dom;
    }
    ~OneLevel() { delete dom; }
    void load_all(const std::string& in) {
        load(in + "/u.bin", dom->u->at, dom->u->size);
        if (dom->uprev != dom->u) load(in + "/uprev.bin", dom->uprev->at, dom->uprev->size);
        load(in + "/f.bin", dom->f->at, dom->f->size);
        load(in + "/r.bin", dom->r->at, dom->r->size);
        Boundary<T>* b[6] = {dom->north, dom->south, dom->east, dom->west, dom->top, dom->bottom};
        const char* nm[6] = {"gN", "gS", "gE", "gW", "gT", "gB"};
        for (int q = 0; q < 6; q++) load(in + "/" + nm[q] + ".bin", b[q]->arr.at, b[q]->arr.size);
        dom->u->to_device();
        if (dom->uprev != dom->u) dom->uprev->to_device();
        dom->f->to_device();
        dom->r->to_device();
        for (int q = 0; q < 6; q++) b[q]->to_device();
    }
    void save_all(const std::string& out) {
        dom->u->to_host();
        dom->r->to_host();
        if (dom->uprev != dom->u) dom->uprev->to_host();
        save(out + "/u.bin", dom->u->at, dom->u->size);
        if (dom->uprev != dom->u) save(out + "/uprev.bin", dom->uprev->at, dom->uprev->size);
        save(out + "/r.bin", dom->r->at, dom->r->size);
    }
};

// Single-operator runs on caller-provided arrays (one level, one device).
template <class T>
int run_op(const Params& p, const std::string& in, const std::string& out) {
    const std::string op = p.s("op");
    if (op == "residual") {
        OneLevel<T> L(p, false);
        L.load_all(in);
        residual<T>(*L.dom, true);
        L.save_all(out);
    } else if (op == "relax_gs") {
        OneLevel<T> L(p, false);
        L.load_all(in);
        GaussSeidel<T> gs;
        static_cast<Relaxation<T>&>(gs).relax(L.levels, (T)p.f("omega"), p.i("nsmooth"), 0, 1);
        L.save_all(out);
    } else if (op == "relax_jacobi") {
        OneLevel<T> L(p, true);
        L.load_all(in);
        Jacobi<T> jac;
        static_cast<Relaxation<T>&>(jac).relax(L.levels, (T)p.f("omega"), p.i("nsmooth"), 0, 1);
        // after an odd number of sweeps u/uprev are swapped; report which
        // buffer "u" now is by saving through the domain's current pointers
        L.save_all(out);
    } else if (op == "gs_kernel") {
        // raw half-sweeps without the boundary refresh (relaxation_kernel only)
        OneLevel<T> L(p, false);
        L.load_all(in);
        GaussSeidel<T> gs;
        const long n = p.i("halfsweeps");
        DeviceArray<T>& u = *L.dom->u;
        for (long s = 0; s < n; s++) {
            gs.relaxation_kernel(*L.dom, (T)p.f("omega"), L.dom->west->is_non_eliminated(),
                                 u.shape[0] - L.dom->east->is_non_eliminated(),
                                 L.dom->south->is_non_eliminated(),
                                 u.shape[1] - L.dom->north->is_non_eliminated(),
                                 L.dom->bottom->is_non_eliminated(),
                                 u.shape[2] - L.dom->top->is_non_eliminated());
        }
        L.save_all(out);
    } else if (op == "restrict_inj" || op == "restrict_fw" || op == "prolong") {
        // uin/uout are standalone halo-padded arrays: shapes and halos given
        Halo hin(p.i("in_hE"), p.i("in_hW"), p.i("in_hN"), p.i("in_hS"), p.i("in_hT"), p.i("in_hB"));
        Halo hout(p.i("out_hE"), p.i("out_hW"), p.i("out_hN"), p.i("out_hS"), p.i("out_hT"), p.i("out_hB"));
        Settings si = make_settings(p);
        si.dims[0] = p.i("in_nx"); si.dims[1] = p.i("in_ny"); si.dims[2] = p.i("in_nz");
        Settings so = si;
        so.dims[0] = p.i("out_nx"); so.dims[1] = p.i("out_ny"); so.dims[2] = p.i("out_nz");
        DeviceArray<T> a(si, hin), b(so, hout);
        load(in + "/in.bin", a.at, a.size);
        load(in + "/out.bin", b.at, b.size);
        a.to_device();
        b.to_device();
        if (op == "restrict_inj") { Injection<T> R; R.restrict_to(a, b); }
        else if (op == "restrict_fw") { FullWeighting<T> R; R.restrict_to(a, b); }
        else { TrilinearInterpolation<T> P; P.prolong(a, b); }
        b.to_host();
        save(out + "/out.bin", b.at, b.size);
    } else if (op == "bc_restrict") {
        // coarse_boundary.restrict_to(u_fine, fine_boundary, fine_settings, R)
        // for every face of a two-level hierarchy
        Settings sf = make_settings(p);
        sf.levels = 2;
        Grid grid(sf, 2);
        BoundaryCondition bc = parse_bc(p.s("bc", "DDDDDD"));
        int_t dev[3] = {0, 0, 0};
        int_t shape[3] = {1, 1, 1};
        Domain<T> fine(grid.domainsettings[0], bc, false, dev, shape);
        Domain<T> coarse(grid.domainsettings[1], bc, false, dev, shape);
        load(in + "/u.bin", fine.u->at, fine.u->size);
        fine.u->to_device();
        Boundary<T>* bf[6] = {fine.north, fine.south, fine.east, fine.west, fine.top, fine.bottom};
        Boundary<T>* bcs[6] = {coarse.north, coarse.south, coarse.east, coarse.west, coarse.top, coarse.bottom};
        const char* nm[6] = {"gN", "gS", "gE", "gW", "gT", "gB"};
        for (int q = 0; q < 6; q++) {
            load(in + "/" + nm[q] + ".bin", bf[q]->arr.at, bf[q]->arr.size);
            load(in + "/c" + nm[q] + ".bin", bcs[q]->arr.at, bcs[q]->arr.size);
            bf[q]->to_device();
            bcs[q]->to_device();
        }
        const bool fw = p.s("restriction") == "fw";
        Injection<T> inj;
        FullWeighting<T> fwr;
        Restriction<T>& R = fw ? (Restriction<T>&)fwr : (Restriction<T>&)inj;
        for (int q = 0; q < 6; q++) bcs[q]->restrict_to(*fine.u, *bf[q], fine.settings, R);
        for (int q = 0; q < 6; q++) {
            bcs[q]->arr.to_host();
            save(out + "/c" + std::string(nm[q]) + ".bin", bcs[q]->arr.at, bcs[q]->arr.size);
        }
    } else {
        fprintf(stderr, "unknown op %s\n", op.c_str());
        return 2;
    }
    return 0;
}

// The reference's VTK writers (include/array.h:148-269) on a fixed field:
// v(i,j,k) = sin(0.1 + 0.37 i + 1.1 j + 2.3 k), shape nx x ny x nz, one halo
// layer east/north/top filled with -(index+1) so print_halo shows it.
int run_vtk(const Params& p, const std::string& out) {
    Settings s;
    s.origin[0] = p.f("ox", 0.5);
    s.origin[1] = p.f("oy", -1.0);
    s.origin[2] = p.f("oz", 2.0);
    s.h = p.f("h", 0.25);
    const uint_t nx = (uint_t)p.i("nx", 3), ny = (uint_t)p.i("ny", 4), nz = (uint_t)p.i("nz", 5);
    Array<double> a(nx, ny, nz);
    for (uint_t i = 0; i < nx; i++)
        for (uint_t j = 0; j < ny; j++)
            for (uint_t k = 0; k < nz; k++) a.at[a.idx(i, j, k)] = std::sin(0.1 + 0.37 * i + 1.1 * j + 2.3 * k);
    a.print(s, (out + "/u.vtk").c_str());
    Halo hl(1, 0, 1, 0, 1, 0);
    Array<double> b(nx, ny, nz, hl);
    for (uint_t t = 0; t < b.size; t++) b.at[t] = -(double)(t + 1);
    for (uint_t i = 0; i < nx; i++)
        for (uint_t j = 0; j < ny; j++)
            for (uint_t k = 0; k < nz; k++) b.at[b.idx(i, j, k)] = std::sin(0.1 + 0.37 * i + 1.1 * j + 2.3 * k);
    b.print_halo(s, (out + "/u_halo.vtk").c_str());
    return 0;
}

// Timing of the reference's own solve on the host cores (the CPU baseline):
// one solve() with maxiter = miniter = ncycles and tolerance 0, so that it
// runs exactly ncycles cycles (each followed by the residual and its norm,
// libpoisson.h:254-353).  The reference prints "<cycle> took <s> seconds"
// per cycle itself; the harness adds the set-up and solve wall times.
template <class T, template <class> class R, template <class> class S>
int run_time(const Params& p) {
    Settings settings = make_settings(p);
    settings.maxiter = p.i("ncycles", 1);
    settings.miniter = settings.maxiter;
    settings.tolerance = 0.0;
    BoundaryCondition bc = parse_bc(p.s("bc", "DDDDDD"));
    const double t0 = omp_get_wtime();
    PoissonSolver<T, R, TrilinearInterpolation, S> solver(settings, bc);
    solver.init();
    solver.to_device();
    const double t1 = omp_get_wtime();
    solver.solve(p.s("cycle", "Fcycle"), p.i("nsmooth", 2), (T)p.f("omega", 1.0));
    const double t2 = omp_get_wtime();
    printf("harness setup_s %.6f solve_s %.6f iterations %.0f rel_res %.17g threads %d\n", t1 - t0, t2 - t1,
           (double)solver.solve_iterations(), (double)solver.relative_residual(), omp_get_max_threads());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s params in-dir out-dir\n", argv[0]);
        return 2;
    }
    Params p = read_params(argv[1]);
    g_problem = (int)p.i("problem", 0);
    const std::string in = argv[2], out = argv[3];
    const bool dbl = p.s("dtype", "f64") == "f64";
    const std::string mode = p.s("mode", "solve");
    if (mode == "op") return dbl ? run_op<double>(p, in, out) : run_op<float>(p, in, out);
    if (mode == "vtk") return run_vtk(p, out);
    const bool fw = p.s("restriction", "inj") == "fw";
    const bool gs = p.s("relaxation", "gs") == "gs";
    if (mode == "time") {
        if (!dbl) return 2;
        if (fw) return gs ? run_time<double, FullWeighting, GaussSeidel>(p) : run_time<double, FullWeighting, Jacobi>(p);
        return gs ? run_time<double, Injection, GaussSeidel>(p) : run_time<double, Injection, Jacobi>(p);
    }
    if (dbl) {
        if (fw) return gs ? run_solve<double, FullWeighting, GaussSeidel>(p, out) : run_solve<double, FullWeighting, Jacobi>(p, out);
        return gs ? run_solve<double, Injection, GaussSeidel>(p, out) : run_solve<double, Injection, Jacobi>(p, out);
    }
    if (fw) return gs ? run_solve<float, FullWeighting, GaussSeidel>(p, out) : run_solve<float, FullWeighting, Jacobi>(p, out);
    return gs ? run_solve<float, Injection, GaussSeidel>(p, out) : run_solve<float, Injection, Jacobi>(p, out);
}
