// The solver-level C entry points: PoissonSolver<T, R, TrilinearInterpolation, S>
// instantiations (T in {double, float}, R in {Injection, FullWeighting},
// S in {Jacobi, GaussSeidel}) behind an opaque handle, for FFI callers.
#include <memory>
#include <stdexcept>
#include <string>

#include "libpoisson.h"
#include "mgp.h"
#include "mgp_problems.h"

namespace {

thread_local std::string g_last_error;

struct SolverBase {
    virtual ~SolverBase() {}
    virtual void init_fn(Poisson::funptr u, Poisson::funptr f, Poisson::funptr dx, Poisson::funptr dy,
                         Poisson::funptr dz) = 0;
    virtual void init_zero() = 0;
    virtual void to_device() = 0;
    virtual void to_device_zero() = 0;
    virtual void to_host() = 0;
    virtual void solve(const std::string& c, int64_t nsmooth, double omega, double maxtime) = 0;
    virtual void cycle(const std::string& c, int64_t nsmooth, double omega) = 0;
    virtual double relres() = 0;
    virtual void sync() = 0;
    virtual void reset() = 0;
    virtual void stats(double* rr, double* t, int64_t* it) = 0;
    virtual int array(int level, int which, void** host, mgp_array* dev) = 0;
    virtual int face(int level, int loc, void** host, mgp_array* dev, int* type) = 0;
    virtual void save(const char* path) = 0;
};

template <class T, template <class> class R, template <class> class S>
struct SolverImpl : SolverBase {
    Poisson::PoissonSolver<T, R, Poisson::TrilinearInterpolation, S> s;
    SolverImpl(Poisson::Settings& st, Poisson::BoundaryCondition& bc) : s(st, bc) {}
    void init_fn(Poisson::funptr u, Poisson::funptr f, Poisson::funptr dx, Poisson::funptr dy,
                 Poisson::funptr dz) override {
        s.init(u, f, dx, dy, dz);
    }
    void init_zero() override { s.init_zero(); }
    void to_device() override { s.to_device(); }
    void to_device_zero() override { s.to_device_zero(); }
    void to_host() override { s.to_host(); }
    void solve(const std::string& c, int64_t nsmooth, double omega, double maxtime) override {
        s.solve(c, nsmooth, (T)omega, "", maxtime);
    }
    void cycle(const std::string& c, int64_t nsmooth, double omega) override { s.cycle(c, nsmooth, (T)omega); }
    double relres() override { return s.relres(); }
    void sync() override { Poisson::mgpx::check(mgp_stream_sync(s.get_stream()), "sync"); }
    void reset() override { s.reset_solution(); }
    void stats(double* rr, double* t, int64_t* it) override {
        if (rr) *rr = (double)s.relative_residual();
        if (t) *t = (double)s.solve_time();
        if (it) *it = (int64_t)s.solve_iterations();
    }
    int array(int level, int which, void** host, mgp_array* dev) override {
        if (level < 0 || level >= (int)s.get_levels()) return MGP_ERR_ARG;
        Poisson::Domain<T>& d = s.get_domain(0, level);
        Poisson::DeviceArray<T>* a = nullptr;
        switch (which) {
        case 'u': a = d.u; break;
        case 'p': a = d.uprev; break;
        case 'f': a = d.f; break;
        case 'r': a = d.r; break;
        default: return MGP_ERR_ARG;
        }
        if (host) *host = a->at;
        if (dev) *dev = a->desc();
        return MGP_OK;
    }
    int face(int level, int loc, void** host, mgp_array* dev, int* type) override {
        if (level < 0 || level >= (int)s.get_levels() || loc < 0 || loc > 5) return MGP_ERR_ARG;
        Poisson::Boundary<T>* b = s.get_domain(0, level).face(loc);
        if (host) *host = b->arr.at;
        if (dev) *dev = b->arr.desc();
        if (type) *type = b->is_non_eliminated() ? MGP_DIRICHLET : MGP_NEUMANN;
        return MGP_OK;
    }
    void save(const char* path) override { s.save(0, path); }
};

// FP32 cycles inside FP64 refinement; host arrays are the FP64 ones
template <template <class> class R, template <class> class S>
struct MixedImpl : SolverBase {
    Poisson::RefinementSolver<R, Poisson::TrilinearInterpolation, S> s;
    MixedImpl(Poisson::Settings& st, Poisson::BoundaryCondition& bc) : s(st, bc) {}
    void init_fn(Poisson::funptr u, Poisson::funptr f, Poisson::funptr dx, Poisson::funptr dy,
                 Poisson::funptr dz) override {
        s.init(u, f, dx, dy, dz);
    }
    void init_zero() override { s.init_zero(); }
    void to_device() override { s.to_device(); }
    void to_device_zero() override { s.to_device_zero(); }
    void to_host() override { s.to_host(); }
    void solve(const std::string& c, int64_t nsmooth, double omega, double maxtime) override {
        s.solve(c, nsmooth, omega, "", maxtime);
    }
    void cycle(const std::string& c, int64_t nsmooth, double omega) override {
        (void)c, (void)nsmooth, (void)omega;
        throw std::invalid_argument("a refinement solver has no bare cycle; use solve");
    }
    double relres() override { return s.fp64().relres(); }
    void sync() override { Poisson::mgpx::check(mgp_stream_sync(s.get_stream()), "sync"); }
    void reset() override { s.reset_solution(); }
    void stats(double* rr, double* t, int64_t* it) override {
        if (rr) *rr = s.relative_residual();
        if (t) *t = s.solve_time();
        if (it) *it = (int64_t)s.solve_iterations();
    }
    int array(int level, int which, void** host, mgp_array* dev) override {
        if (level != 0) return MGP_ERR_ARG;  // the FP64 problem lives on the finest level only
        Poisson::Domain<double>& d = s.fp64().get_domain(0, 0);
        Poisson::DeviceArray<double>* a = nullptr;
        if (which == 'l') {  // FP32 lo word of the hi + lo iterate (device only)
            const mgp_array lo = s.solution_lo();
            if (!lo.data) return MGP_ERR_ARG;
            if (host) *host = nullptr;
            if (dev) *dev = lo;
            return MGP_OK;
        }
        switch (which) {
        case 'u': a = d.u; break;
        case 'p': a = d.uprev; break;
        case 'f': a = d.f; break;
        case 'r': a = d.r; break;
        default: return MGP_ERR_ARG;
        }
        if (host) *host = a->at;
        if (dev) *dev = a->desc();
        return MGP_OK;
    }
    int face(int level, int loc, void** host, mgp_array* dev, int* type) override {
        if (level != 0 || loc < 0 || loc > 5) return MGP_ERR_ARG;
        Poisson::Boundary<double>* b = s.fp64().get_domain(0, 0).face(loc);
        if (host) *host = b->arr.at;
        if (dev) *dev = b->arr.desc();
        if (type) *type = b->is_non_eliminated() ? MGP_DIRICHLET : MGP_NEUMANN;
        return MGP_OK;
    }
    void save(const char* path) override { s.fp64().save(0, path); }
};

template <class T>
SolverBase* make_t(const mgp_solver_config* c, Poisson::Settings& st, Poisson::BoundaryCondition& bc) {
    const bool fw = c->restriction == MGP_FULL_WEIGHTING;
    const bool gs = c->relaxation == MGP_GAUSS_SEIDEL;
    if (fw) {
        if (gs) return new SolverImpl<T, Poisson::FullWeighting, Poisson::GaussSeidel>(st, bc);
        return new SolverImpl<T, Poisson::FullWeighting, Poisson::Jacobi>(st, bc);
    }
    if (gs) return new SolverImpl<T, Poisson::Injection, Poisson::GaussSeidel>(st, bc);
    return new SolverImpl<T, Poisson::Injection, Poisson::Jacobi>(st, bc);
}

template <class F>
int guarded(F&& fn) {
    try {
        fn();
        return MGP_OK;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return MGP_ERR_ARG;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return MGP_ERR_STATE;
    }
}

SolverBase* S(void* h) { return static_cast<SolverBase*>(h); }

}  // namespace

namespace mgp {
void set_last_error(const char* msg) { g_last_error = msg; }
}  // namespace mgp

extern "C" {

int mgp_solver_create(const mgp_solver_config* c, void** out) {
    if (!c || !out) return MGP_ERR_ARG;
    *out = nullptr;
    for (int q = 0; q < 6; q++)
        if (c->bc[q] != MGP_DIRICHLET && c->bc[q] != MGP_NEUMANN) return MGP_ERR_ARG;
    if ((c->dtype != MGP_F64 && c->dtype != MGP_F32 && c->dtype != MGP_MIXED) ||
        (c->restriction != MGP_INJECTION && c->restriction != MGP_FULL_WEIGHTING) ||
        (c->relaxation != MGP_JACOBI && c->relaxation != MGP_GAUSS_SEIDEL) ||
        (c->mode != MGP_MODE_REFERENCE && c->mode != MGP_MODE_CONSISTENT))
        return MGP_ERR_ARG;
    return guarded([&] {
        Poisson::Settings st;
        for (int d = 0; d < 3; d++) {
            if (c->dims[d] < 3) throw std::invalid_argument("dims must be >= 3");
            st.dims[d] = c->dims[d];
            st.origin[d] = c->origin[d];
        }
        st.levels = c->levels;
        st.h = c->h;
        st.lengthx = c->h * (double)(c->dims[0] - 1);
        st.maxiter = c->maxiter;
        st.miniter = c->miniter;
        st.tolerance = c->tolerance;
        st.coarse_sweeps = c->coarse_sweeps;
        st.mode = c->mode;
        st.use_graphs = !(c->flags & MGP_SOLVER_NO_GRAPHS);
        st.fused_sweeps = !(c->flags & MGP_SOLVER_NO_FUSED);
        st.print_cycle_time = !(c->flags & MGP_SOLVER_QUIET);
        st.print_settings = !(c->flags & MGP_SOLVER_QUIET);
        st.refine_hilo = !(c->flags & MGP_SOLVER_PLAIN_ITERATE);
        st.refine_coarse_fp64 = !(c->flags & MGP_SOLVER_COARSE_FP32);
        st.coarse_cg = (c->flags & MGP_SOLVER_COARSE_CG) != 0;
        if (st.coarse_cg && c->mode != MGP_MODE_CONSISTENT)
            throw std::invalid_argument("MGP_SOLVER_COARSE_CG needs MGP_MODE_CONSISTENT");
        if (c->refine_polish < 0) throw std::invalid_argument("refine_polish must be >= 0");
        st.refine_polish = c->refine_polish;
        if (c->half_from < 0) throw std::invalid_argument("half_from must be >= 0");
        if (c->half_from > 0 && c->dtype != MGP_MIXED)
            throw std::invalid_argument("half_from needs dtype MGP_MIXED");
        st.half_from = (Poisson::uint_t)c->half_from;
        auto t = [&](int q) { return c->bc[q] == MGP_NEUMANN ? Poisson::NEUMANN : Poisson::DIRICHLET; };
        Poisson::BoundaryCondition bc{t(MGP_NORTH), t(MGP_SOUTH), t(MGP_EAST), t(MGP_WEST), t(MGP_TOP), t(MGP_BOTTOM)};
        if (c->dtype == MGP_MIXED) {
            const bool fw = c->restriction == MGP_FULL_WEIGHTING, gs = c->relaxation == MGP_GAUSS_SEIDEL;
            if (fw) *out = gs ? (SolverBase*)new MixedImpl<Poisson::FullWeighting, Poisson::GaussSeidel>(st, bc)
                              : (SolverBase*)new MixedImpl<Poisson::FullWeighting, Poisson::Jacobi>(st, bc);
            else *out = gs ? (SolverBase*)new MixedImpl<Poisson::Injection, Poisson::GaussSeidel>(st, bc)
                           : (SolverBase*)new MixedImpl<Poisson::Injection, Poisson::Jacobi>(st, bc);
        } else {
            *out = c->dtype == MGP_F64 ? make_t<double>(c, st, bc) : make_t<float>(c, st, bc);
        }
    });
}

int mgp_solver_destroy(void* h) {
    return guarded([&] { delete S(h); });
}

int mgp_solver_init_problem(void* h, int problem) {
    if (!h || problem < 0 || problem > 3) return MGP_ERR_ARG;
    return guarded([&] {
        mgp::problems::current() = problem;
        S(h)->init_fn(&mgp::problems::u, &mgp::problems::f, &mgp::problems::dudx, &mgp::problems::dudy,
                      &mgp::problems::dudz);
    });
}

int mgp_solver_init_fn(void* h, double (*u)(double, double, double), double (*f)(double, double, double),
                       double (*dx)(double, double, double), double (*dy)(double, double, double),
                       double (*dz)(double, double, double)) {
    if (!h || !u || !f || !dx || !dy || !dz) return MGP_ERR_ARG;
    return guarded([&] { S(h)->init_fn(u, f, dx, dy, dz); });
}

int mgp_solver_init_zero(void* h) { return h ? guarded([&] { S(h)->init_zero(); }) : MGP_ERR_ARG; }
int mgp_solver_to_device(void* h) { return h ? guarded([&] { S(h)->to_device(); }) : MGP_ERR_ARG; }
int mgp_solver_to_device_zero(void* h) { return h ? guarded([&] { S(h)->to_device_zero(); }) : MGP_ERR_ARG; }
int mgp_solver_to_host(void* h) { return h ? guarded([&] { S(h)->to_host(); }) : MGP_ERR_ARG; }

int mgp_solver_solve(void* h, const char* cycle, int64_t nsmooth, double omega, double maxtime) {
    if (!h || !cycle) return MGP_ERR_ARG;
    return guarded([&] { S(h)->solve(cycle, nsmooth, omega, maxtime); });
}

int mgp_solver_cycle(void* h, const char* cycle, int64_t nsmooth, double omega) {
    if (!h || !cycle) return MGP_ERR_ARG;
    return guarded([&] { S(h)->cycle(cycle, nsmooth, omega); });
}

int mgp_solver_relres(void* h, double* rr) {
    if (!h || !rr) return MGP_ERR_ARG;
    return guarded([&] { *rr = S(h)->relres(); });
}

int mgp_solver_sync(void* h) { return h ? guarded([&] { S(h)->sync(); }) : MGP_ERR_ARG; }
int mgp_solver_reset(void* h) { return h ? guarded([&] { S(h)->reset(); }) : MGP_ERR_ARG; }

int mgp_solver_stats(void* h, double* rr, double* t, int64_t* it) {
    if (!h) return MGP_ERR_ARG;
    return guarded([&] { S(h)->stats(rr, t, it); });
}

int mgp_solver_array(void* h, int level, int which, void** host, mgp_array* dev) {
    if (!h) return MGP_ERR_ARG;
    return S(h)->array(level, which, host, dev);
}

int mgp_solver_face(void* h, int level, int loc, void** host, mgp_array* dev, int* type) {
    if (!h) return MGP_ERR_ARG;
    return S(h)->face(level, loc, host, dev, type);
}

int mgp_solver_save(void* h, const char* path) {
    if (!h || !path) return MGP_ERR_ARG;
    return guarded([&] { S(h)->save(path); });
}

const char* mgp_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
