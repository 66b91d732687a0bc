// FP64 iterative refinement around FP32 multigrid cycles (addition; the
// reference's PoissonSolver computes in one type T throughout).
//
// RefinementSolver<R, P, S> keeps the problem (f and the boundary data) and the
// solution u in FP64 on the finest level and repeats
//     r  = f - L_h u                    (FP64 residual with the real boundary data)
//     solve L_h e = r approximately     (one FP32 V- or F-cycle from e = 0,
//                                         homogeneous boundary data, all levels)
//     u += e                            (FP64 update)
// until ||r||inf / ||f||inf <= tolerance.  FP32 carries the smoothing,
// restriction and prolongation; FP64 the residual, the update, the
// convergence test and (Settings::refine_coarse_fp64) the arithmetic of the
// coarsest level's smoothing, so the attainable accuracy is that of the FP64
// residual.  It requires SolverMode CONSISTENT (homogeneous coarse Neumann
// data).
//
// Settings::refine_hilo (default) holds the iterate as FP64 hi plus an FP32
// lo word (the update renormalised by an error-free sum, the residual taken
// of hi + lo: mgp_update_residual_hilo_f64_f32).  A plain FP64 u floors
// ||r||/||f|| near 12/h^2 ulp(u)/||f|| (about 1e-10 at h = 1/1024); hi + lo
// carries about 77 bits, so the FP64 residual keeps falling.  get_solution()
// is hi, the FP64 rounding of the iterate; its residual and relative_residual()
// are those of hi + lo.
//
// Settings::half_from > 0 keeps levels half_from and coarser of the FP32
// cycle in FP16 storage (Fp16Levels, fp16levels.h): the FP32 solver then has
// half_from levels and its last one hands the cycle on to the FP16 levels.
#ifndef POISSON_REFINEMENT
#define POISSON_REFINEMENT

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "fp16levels.h"

namespace Poisson {

template <class T, template <class> class R, template <class> class P, template <class> class S>
class PoissonSolver;

template <template <class> class R, template <class> class P, template <class> class S>
class RefinementSolver {
public:
    RefinementSolver(Settings& s, BoundaryCondition& bc)
        : settings(s), outer_settings(s), inner_settings(s), BC(bc) {
        if (s.mode != CONSISTENT)
            throw std::invalid_argument("RefinementSolver needs Settings::mode = CONSISTENT");
        outer_settings.levels = 1;
        if (s.half_from > 0) {
            if (s.half_from < 2 || s.half_from >= s.levels)
                throw std::invalid_argument("RefinementSolver: half_from must lie in [2, levels)");
            const bool dirichlet = bc.north == DIRICHLET && bc.south == DIRICHLET && bc.east == DIRICHLET &&
                                   bc.west == DIRICHLET && bc.top == DIRICHLET && bc.bottom == DIRICHLET;
            if (!dirichlet || !std::is_same<S<float>, GaussSeidel<float>>::value ||
                !std::is_same<R<float>, FullWeighting<float>>::value)
                throw std::invalid_argument(
                    "RefinementSolver: FP16 levels need Gauss-Seidel, full weighting and Dirichlet faces");
            inner_settings.levels = s.half_from;
        }
        inner_settings.print_cycle_time = false;
        outer_settings.print_cycle_time = false;
        inner_settings.coarse_fp64 = s.refine_coarse_fp64;
        inner_settings.print_settings = false;  // the FP64 levels print, as one solver would
        outer = new PoissonSolver<double, R, P, S>(outer_settings, BC);
        try {
            inner = new PoissonSolver<float, R, P, S>(inner_settings, BC);
        } catch (...) {
            delete outer;
            throw;
        }
        inner->use_stream(outer->get_stream());
        if (s.half_from > 0) {
            try {
                mgpx::StreamScope scope(outer->get_stream());
                Domain<float>& last = inner->get_domain(0, s.half_from - 1);
                half = new Fp16Levels(last, s.levels - s.half_from, s.coarse_sweeps);
                last.below = half;
            } catch (...) {
                delete inner;
                delete outer;
                throw;
            }
        }
        void* p = nullptr;
        mgpx::check(mgpx::dev_alloc(&p, 8 * sizeof(double)), "RefinementSolver");
        dev_norm = static_cast<double*>(p);
        mgpx::check(mgpx::host_alloc(&p, 8 * sizeof(double)), "RefinementSolver");
        host_norm = static_cast<double*>(p);
        if (s.refine_hilo) alloc_hilo();
    }
    RefinementSolver(const RefinementSolver&) = delete;
    RefinementSolver& operator=(const RefinementSolver&) = delete;
    ~RefinementSolver() {
        // the FP32 levels' arrays are in use by work queued on the FP64
        // solver's stream (not waited for here when the stream is adopted)
        if (outer->owns_its_stream()) mgp_stream_sync(outer->get_stream());
        delete inner;
        delete half;
        delete outer;
        mgpx::dev_free(dev_norm);
        mgpx::host_free(host_norm);
        for (void* q : lo) mgpx::dev_free(q);
        mgpx::dev_free(own_alt);
    }

    void init() { this->init(&ufun, &ffun, &dudxfun, &dudyfun, &dudzfun); }
    void init(funptr u, funptr f, funptr dx, funptr dy, funptr dz) {
        outer->init(u, f, dx, dy, dz);
        inner->init_zero();
    }
    void init_zero() {
        outer->init_zero();
        inner->init_zero();
    }
    void to_device() {
        outer->to_device();
        inner->init_zero();
        inner->to_device();
    }
    void to_host() { outer->to_host(); }
    // to_device() of a freshly constructed (all-zero) solver, made on the device
    void to_device_zero() {
        outer->to_device_zero();
        inner->to_device_zero();
    }
    void verbose(bool onoff) { is_verbose = onoff; }
    void reset_solution() { outer->reset_solution(); }

    void solve(const std::string solver = "Fcycle", const int_t nsmooth = 2, double omega = -1.0,
               std::string ofile = "", double maxtime = 20 * 60) {
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        auto since = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
        (void)ofile;
        start(solver);
        if (is_verbose) std::cout << std::setw(4) << 0 << ": Initial residual: " << std::setw(8) << rel_res << std::endl;
        if (settings.maxiter > 0) {
            while (true) {
                enqueue_iteration(nsmooth, omega);
                mgpx::check(mgp_stream_sync(outer->get_stream()), "refinement");
                if (is_verbose)
                    std::cout << std::setw(4) << iter + 1 << ": Relative residual: " << std::setw(8)
                              << host_norm[0] / fnorm << std::endl;
                if (iteration_done(since() > maxtime)) break;
            }
            // the iterations wrote r only as FP32: the FP64 r of the final u
            // (of hi + lo: the residual the stop rule saw)
            mgpx::StreamScope scope(outer->get_stream());
            final_residual();
            mgpx::check(mgp_stream_sync(outer->get_stream()), "refinement");
        }
        wtime = since();
    }

    // ---- the same iteration split for callers that drive many solves
    // without blocking (one stream each): start() (waits: ||f|| and the
    // initial residual), then per iteration enqueue_iteration() and, once
    // the stream has passed it (event after it), iteration_done(), which
    // applies the stop rule of solve() and returns true when the solve ends.
    void start(const std::string& solver) {
        const bool v = solver == "vcycle" || solver == "Vcycle";
        const bool f = solver == "fcycle" || solver == "Fcycle";
        if (!(v || f)) throw std::invalid_argument("RefinementSolver: solver must be \"Vcycle\" or \"Fcycle\"");
        fcycle = f;
        void* st = outer->get_stream();
        mgpx::StreamScope scope(st);
        Domain<double>& d64 = outer->get_domain(0, 0);
        fnorm = d64.f->infinity_norm();
        // r = 0 first: the residual kernels write the box points only, and the
        // FP32 right-hand side is cast from all of r, boundary nodes included,
        // which the F-cycle's restriction of f reads (a caller may have used
        // r as scratch between solves, as the server does for results)
        d64.r->init_zero();
        // a solve starts from the FP64 u as stored: lo = 0
        if (lo[0]) {
            const mgp_array l = lo_desc(cur);
            mgpx::check(mgp_fill_f32(&l, 0.0, st), "refinement: lo");
        }
        enqueue_norm(d64, st);
        mgpx::check(mgp_stream_sync(st), "refinement: norm");
        rel_res = host_norm[0] / fnorm;
        best = rel_res;
        stall = 0;
        iter = 0;
        f32_ready = false;
    }
    // f32 = r, e = 0, one FP32 cycle, u += e, then the FP64 residual and its
    // norm (copied to host memory the stream writes when it gets there).  The
    // residual goes straight to the FP32 right-hand side of the next
    // iteration ((float) of the same FP64 value a separate cast would read).
    void enqueue_iteration(int_t nsmooth, double omega) {
        void* st = outer->get_stream();
        mgpx::StreamScope scope(st);
        Domain<double>& d64 = outer->get_domain(0, 0);
        Domain<float>& d32 = inner->get_domain(0, 0);
        const mgp_array r64 = d64.r->desc(), f32 = d32.f->desc(), u64 = d64.u->desc();
        if (!f32_ready) mgpx::check(mgp_cast_f64_f32(&r64, &f32, st), "refinement: cast");
        // e = 0: taken as zero by the cycle's first smoothing, which then does
        // not read it (written as zeros wherever it would be read instead)
        d32.u_zero = true;
        inner->cycle(fcycle ? "Fcycle" : "Vcycle", nsmooth, (float)omega);
        d32.materialize_zero();  // (consumed by every cycle; kept for any path that does not)
        const mgp_array e32 = d32.u->desc();
        mgpx::check(mgp_zero_scalars(dev_norm, 1, st), "refinement: norm");
        if (lo[0]) {
            // hi + lo iterate: the update renormalised into the second pair
            // of buffers and the residual of it, in one pass
            int64_t b[6];
            d64.box(b);
            const mgp_array f64 = d64.f->desc(), r64v = d64.r->desc(), l0 = lo_desc(cur), l1 = lo_desc(cur ^ 1);
            mgp_array un = u64;
            un.data = hi_alt();
            const mgp_faces F = d64.faces();
            mgpx::check(mgp_update_residual_hilo_f64_f32(&u64, &un, &l0, &l1, &e32, &f64, &r64v, &f32, &F, b, dev_norm,
                                                         st),
                        "refinement: update + residual");
            swap_hi();
            cur ^= 1;
        } else if (settings.refine_polish > 0 || !update_residual_norm_to_f32(d64, e32, f32, dev_norm)) {
            // update and residual in one pass where nothing runs between them
            mgpx::check(mgp_add_f32_to_f64(&u64, &e32, st), "refinement: update");
            // optional FP64 smoothing of the updated solution: settles u at a
            // fixed point of the FP64 smoother, whose residual sits below the
            // rounding floor of an unsmoothed FP64 u at fine h
            if (settings.refine_polish > 0) outer->smooth_finest(settings.refine_polish, omega);
            residual_norm_to_f32(outer->get_domain(0, 0), f32, dev_norm);
        }
        mgpx::check(mgp_memcpy_d2h(host_norm, dev_norm, sizeof(double), st), "refinement: norm");
        f32_ready = true;
    }
    // stop rule of solve(); `timed_out` stands for the wall-time limit.
    // Consistent-mode rule: tolerance reached, divergence (norm > 2x the
    // previous), wall time, maxiter, or stagnation: no new smallest norm in
    // the last `stall_limit` iterations (a floor the iteration cannot pass;
    // the solve then ends with the tolerance not reached)
    static constexpr int stall_limit = 4;
    bool iteration_done(bool timed_out = false) {
        const double norm = host_norm[0] / fnorm;
        if (norm <= settings.tolerance) {
            rel_res = norm;
            return true;
        }
        if (norm < best) {
            best = norm;
            stall = 0;
        } else {
            stall++;
        }
        if (iter >= settings.miniter && (norm > 2.0 * rel_res || timed_out || stall >= stall_limit)) {
            rel_res = norm;
            return true;
        }
        rel_res = norm;
        iter++;
        return iter >= settings.maxiter;
    }
    // the final residual of a solve driven through start()/enqueue_iteration()
    // (solve() calls it itself): the FP64 r of the final iterate, enqueued
    void final_residual() {
        Domain<double>& d64 = outer->get_domain(0, 0);
        if (!lo[0]) {
            residual<double>(d64, true);
            return;
        }
        int64_t b[6];
        d64.box(b);
        const mgp_array u = d64.u->desc(), f = d64.f->desc(), r = d64.r->desc(), l = lo_desc(cur);
        const mgp_faces F = d64.faces();
        mgpx::check(mgp_residual_hilo_f64(&u, &l, &f, &r, &F, b, nullptr, mgpx::stream()), "refinement: residual");
    }
    // device descriptor of the FP32 lo word of the iterate (level 0's layout);
    // data NULL where the iterate is plain FP64
    mgp_array solution_lo() const {
        if (lo[0]) return lo_desc(cur);
        mgp_array a = outer->get_domain(0, 0).u->desc();
        a.data = nullptr;
        return a;
    }
    double relative_residual() { return rel_res; }
    double solve_time() { return wtime; }
    double solve_iterations() { return (double)iter; }
    DeviceArray<double>& get_rhs(uint_t g) { return outer->get_rhs(g); }
    const DeviceArray<double>& get_solution(uint_t g) { return outer->get_solution(g); }
    const DeviceArray<double>& get_residual(uint_t g) { return outer->get_residual(g); }
    Boundary<double>& get_top_bc(uint_t g) { return outer->get_top_bc(g); }
    Boundary<double>& get_bottom_bc(uint_t g) { return outer->get_bottom_bc(g); }
    PoissonSolver<double, R, P, S>& fp64() { return *outer; }
    PoissonSolver<float, R, P, S>& fp32() { return *inner; }
    Fp16Levels* fp16() { return half; }  // nullptr without FP16 levels
    void* get_stream() const { return outer->get_stream(); }
    // enqueue on `s` (owned by the caller) from now on (waits for the current
    // stream unless wait = false; recorded cycle graphs are dropped)
    void use_stream(void* s, bool wait = true) {
        outer->use_stream(s, wait);
        inner->use_stream(s, false);
    }

private:
    void enqueue_norm(Domain<double>& d, void* st) {
        mgpx::check(mgp_zero_scalars(dev_norm, 1, st), "refinement: norm");
        if (lo[0]) {
            int64_t b[6];
            d.box(b);
            const mgp_array u = d.u->desc(), f = d.f->desc(), r = d.r->desc(), l = lo_desc(cur);
            const mgp_faces F = d.faces();
            mgpx::check(mgp_residual_hilo_f64(&u, &l, &f, &r, &F, b, dev_norm, st), "refinement: residual");
        } else {
            residual_norm<double>(d, dev_norm);
        }
        mgpx::check(mgp_memcpy_d2h(host_norm, dev_norm, sizeof(double), st), "refinement: norm");
    }
    // the FP32 lo words (two buffers: the update writes the other one) and,
    // where the FP64 level has no second u buffer, one of its own
    void alloc_hilo() {
        Domain<double>& d64 = outer->get_domain(0, 0);
        const size_t n = (size_t)d64.u->size;
        mgpx::StreamScope scope(outer->get_stream());
        for (int k = 0; k < 2; k++) {
            mgpx::check(mgpx::dev_alloc(&lo[k], n * sizeof(float)), "RefinementSolver: lo");
            const mgp_array l = lo_desc(k);
            mgpx::check(mgp_fill_f32(&l, 0.0, outer->get_stream()), "RefinementSolver: lo");
        }
        if (!d64.u_alt) {
            void* p = nullptr;
            mgpx::check(mgpx::dev_alloc(&p, n * sizeof(double)), "RefinementSolver: u");
            own_alt = static_cast<double*>(p);
        }
        mgpx::check(mgp_stream_sync(outer->get_stream()), "RefinementSolver: lo");
    }
    mgp_array lo_desc(int k) const {
        mgp_array a = outer->get_domain(0, 0).u->desc();
        a.data = lo[k];
        return a;
    }
    void* hi_alt() { return own_alt ? (void*)own_alt : (void*)outer->get_domain(0, 0).u_alt; }
    void swap_hi() {
        Domain<double>& d64 = outer->get_domain(0, 0);
        if (own_alt) {
            std::swap(d64.u->devptr, own_alt);
        } else {
            d64.swap_u_alt();
        }
    }

    Settings settings, outer_settings, inner_settings;
    BoundaryCondition BC;
    PoissonSolver<double, R, P, S>* outer = nullptr;
    PoissonSolver<float, R, P, S>* inner = nullptr;
    Fp16Levels* half = nullptr;
    double* dev_norm = nullptr;
    double* host_norm = nullptr;
    bool is_verbose = false;
    bool fcycle = true;
    bool f32_ready = false;  // the FP32 right-hand side already holds (float) r
    void* lo[2] = {nullptr, nullptr};  // FP32 lo word of the hi + lo iterate (ping-pong)
    double* own_alt = nullptr;         // second hi buffer where level 0 has no u_alt
    int cur = 0;                       // lo[cur] belongs to the current u
    uint_t iter = 0;
    int stall = 0;
    double rel_res = 0.0, wtime = 0.0, fnorm = 1.0, best = 0.0;
};
}  // namespace Poisson
#endif
