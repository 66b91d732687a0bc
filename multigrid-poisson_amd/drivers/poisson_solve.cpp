// Example driver over the drop-in API (the reference's drivers/dirichlet.cpp
// pattern): build a PoissonSolver, init() an analytic problem, to_device(),
// solve(), to_host(), report the max-abs error against the analytic solution
// and the reference's stats line.
//
//   poisson_solve -x 65 -y 65 -z 65 -l 4 -length 1 -maxiter 10 -tol 0
//                 [-problem 0..3] [-bc DDDDDD] [-cycle Vcycle|Fcycle]
//                 [-relax gs|jacobi] [-restrict inj|fw] [-nsmooth 2] [-omega 1]
//                 [-float] [-consistent] [-nographs] [-stats file]
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "libpoisson.h"
#include "mgp_problems.h"
#include "parser.h"

namespace Poisson {
double_t ufun(double_t x, double_t y, double_t z) { return mgp::problems::u(x, y, z); }
double_t ffun(double_t x, double_t y, double_t z) { return mgp::problems::f(x, y, z); }
double_t dudxfun(double_t x, double_t y, double_t z) { return mgp::problems::dudx(x, y, z); }
double_t dudyfun(double_t x, double_t y, double_t z) { return mgp::problems::dudy(x, y, z); }
double_t dudzfun(double_t x, double_t y, double_t z) { return mgp::problems::dudz(x, y, z); }
}  // namespace Poisson

using namespace Poisson;

struct Opts {
    std::string bc = "DDDDDD", cycle = "Vcycle", relax = "gs", restr = "inj";
    int nsmooth = 2;
    double omega = -1.0;
    bool single = false;
};

template <class T, template <class> class R, template <class> class S>
int run(Settings& settings, BoundaryCondition& bc, const Opts& o) {
    PoissonSolver<T, R, TrilinearInterpolation, S> solver(settings, bc);
    solver.init();
    solver.verbose(true);
    solver.to_device();
    solver.solve(o.cycle, o.nsmooth, (T)o.omega);
    std::cout << "It took " << solver.solve_time() << " seconds to run " << solver.solve_iterations() << " cycles"
              << std::endl;
    solver.to_host();
    const DeviceArray<T>& u = solver.get_solution(0);
    double maxerr = 0.0;
    for (uint_t i = 0; i < u.shape[0]; i++)
        for (uint_t j = 0; j < u.shape[1]; j++)
            for (uint_t k = 0; k < u.shape[2]; k++) {
                const double ut = ufun(settings.origin[0] + ((double_t)i) * settings.h,
                                       settings.origin[1] + ((double_t)j) * settings.h,
                                       settings.origin[2] + ((double_t)k) * settings.h);
                maxerr = std::max(maxerr, std::abs((double)u.at[u.idx(i, j, k)] - ut));
            }
    std::cout << "Maximal error: " << std::setw(8) << maxerr << std::endl;
    if (settings.print_result) solver.save_all("u.vtk", "f.vtk", "r.vtk");
    if (settings.write_final_stats) {  // drivers/dirichlet.cpp:104-118
        std::ofstream out(settings.stats_file, std::ios::app);
        out << "#    seconds     abs_err     rel_res domain_size     spacing     maxiter     lengthx      levels     threads     devices"
            << std::endl;
        out << std::setw(12) << solver.solve_time() << std::setw(12) << maxerr << std::setw(12)
            << solver.relative_residual() << std::setw(12) << settings.dims[0] * settings.dims[1] * settings.dims[2]
            << std::setw(12) << settings.h << std::setw(12) << solver.solve_iterations() << std::setw(12)
            << settings.lengthx << std::setw(12) << settings.levels << std::setw(12) << 1 << std::setw(12) << 1
            << std::endl;
    }
    return 0;
}

int main(int argc, char* argv[]) {
    // split off this driver's own flags; the rest go to the reference's parser
    Opts o;
    int problem = 0;
    bool consistent = false, nographs = false;
    std::vector<char*> rest = {argv[0]};
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto val = [&]() { return std::string(argv[++i]); };
        if (a == "-problem" && i + 1 < argc) problem = std::stoi(val());
        else if (a == "-bc" && i + 1 < argc) o.bc = val();
        else if (a == "-cycle" && i + 1 < argc) o.cycle = val();
        else if (a == "-relax" && i + 1 < argc) o.relax = val();
        else if (a == "-restrict" && i + 1 < argc) o.restr = val();
        else if (a == "-nsmooth" && i + 1 < argc) o.nsmooth = std::stoi(val());
        else if (a == "-omega" && i + 1 < argc) o.omega = std::stod(val());
        else if (a == "-float") o.single = true;
        else if (a == "-consistent") consistent = true;
        else if (a == "-nographs") nographs = true;
        else rest.push_back(argv[i]);
    }
    mgp::problems::current() = problem;
    Settings settings = parser((int)rest.size(), rest.data());
    settings.origin[0] = settings.origin[1] = settings.origin[2] = 0.0;
    if (consistent) settings.mode = CONSISTENT;
    if (nographs) settings.use_graphs = false;
    if (o.bc.size() != 6) {
        std::cerr << "-bc needs six letters D/N in the order north south east west top bottom" << std::endl;
        return 2;
    }
    auto t = [&](int q) { return o.bc[q] == 'N' ? NEUMANN : DIRICHLET; };
    BoundaryCondition bc{t(0), t(1), t(2), t(3), t(4), t(5)};
    const bool fw = o.restr == "fw", gs = o.relax == "gs";
    if (o.single) {
        if (fw) return gs ? run<float, FullWeighting, GaussSeidel>(settings, bc, o) : run<float, FullWeighting, Jacobi>(settings, bc, o);
        return gs ? run<float, Injection, GaussSeidel>(settings, bc, o) : run<float, Injection, Jacobi>(settings, bc, o);
    }
    if (fw) return gs ? run<double, FullWeighting, GaussSeidel>(settings, bc, o) : run<double, FullWeighting, Jacobi>(settings, bc, o);
    return gs ? run<double, Injection, GaussSeidel>(settings, bc, o) : run<double, Injection, Jacobi>(settings, bc, o);
}
