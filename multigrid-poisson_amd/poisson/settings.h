// Solver settings (reference include/settings.h:14-39, src/settings.cpp).
//
// The reference's public fields and defaults are kept.  Fields after
// `stats_file` are this library's additions; their defaults reproduce the
// reference's behaviour.
#ifndef INPUT_SETTINGS
#define INPUT_SETTINGS

#include <iostream>
#include <string>

#include "definitions.h"

#ifdef _OPENMP
#include <omp.h>  // the reference's settings.h includes it; its drivers rely on that
#endif

namespace Poisson {

// How the solve treats boundaries, colours and arithmetic.
//  REFERENCE:  the reference's behaviour exactly, quirks included (north,
//              south and bottom faces are always Dirichlet, coarse Neumann data
//              from one-sided differences, linear-index red/black colouring,
//              stagnation stop rule).
//  CONSISTENT: all six requested faces honoured, homogeneous coarse Neumann
//              correction, (i+j+k) colouring, cancellation-aware stencils and a
//              ||r||/||f|| <= tolerance stop.
enum SolverMode { REFERENCE = 0, CONSISTENT = 1 };

class Settings {
public:
    uint_t levels = 1;
    uint_t dims[3] = {3, 3, 3};
    uint_t maxiter = 100;
    uint_t miniter = 1;
    double_t tolerance = 1e-6;
    int_t host = -1;  // the host (omp_get_initial_device() in the reference)
    int_t dev = 0;    // GPU ordinal
    double_t lengthx = 3.1415;  // length in x
    double_t h = 0.33;
#ifndef PROBLEM2
    double_t origin[3] = {-0.4, 0.21, 1.23};
#else
    double_t origin[3] = {-1, -1, -1};
#endif
    bool print_result = false;
    bool write_final_stats = false;
    std::string stats_file = "";

    // ---- additions
    int mode = REFERENCE;          // SolverMode
    bool use_graphs = true;        // replay each cycle as one captured HIP graph
    bool fused_sweeps = true;      // one-pass red-black sweeps where the colouring allows
    bool print_cycle_time = true;  // the reference prints "<cycle> took <s> seconds" per cycle
    bool print_settings = true;    // the reference prints every level's DomainSettings at construction
    int_t coarse_sweeps = 0;       // CONSISTENT only: smoothing passes per visit of the coarsest
                                   // level (0 = nsmooth, the reference's schedule)
    int_t refine_polish = 0;       // RefinementSolver: FP64 smoothing passes on the finest level
                                   // after each update (0 = none)
    uint_t half_from = 0;          // RefinementSolver: levels >= half_from of the FP32 cycle stored
                                   // in FP16 (0 = none; Gauss-Seidel, full weighting, all faces
                                   // Dirichlet)
    bool refine_hilo = true;       // RefinementSolver: the iterate held as FP64 hi + FP32 lo, so
                                   // that its FP64 residual is not floored by u's rounding to FP64
    bool refine_coarse_fp64 = true;  // RefinementSolver: the FP32 cycle's coarsest level smoothed
                                     // with FP64 arithmetic (sets coarse_fp64 of its FP32 levels)
    bool coarse_fp64 = false;      // FP32 levels: the coarsest level's Gauss-Seidel passes in FP64
                                   // arithmetic where the level fits one workgroup's LDS
    bool coarse_cg = false;        // CONSISTENT only: each visit of the coarsest level solves it by
                                   // CG in FP64 (to 1e-8 relative) where it fits one workgroup,
                                   // instead of coarse_sweeps smoothing passes

    Settings() {}
    Settings(const Settings& s) = default;
    Settings& operator=(const Settings& s) = default;
    ~Settings() {}
};

inline std::ostream& operator<<(std::ostream& os, const Settings& settings) {
    os << "Settings:" << std::endl;
    os << "\tDomain size: (" << settings.dims[0] << "," << settings.dims[1] << "," << settings.dims[2] << ")"
       << std::endl;
    os << "\tMax iterations: " << settings.maxiter << std::endl;
    os << "\tTolerance " << settings.tolerance << std::endl;
    os << "\tHost id: " << settings.host << std::endl;
    os << "\tDevice id: " << settings.dev << std::endl;
    os << "\tSave result: " << settings.print_result << std::endl;
    return os;
}
}  // namespace Poisson
#endif
