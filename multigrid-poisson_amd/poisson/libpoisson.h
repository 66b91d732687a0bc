// The solver façade (reference include/libpoisson.h:28-420):
// PoissonSolver<T, Restriction, Prolongation, Relaxation> over one GPU.
//
// Same constructors, methods, defaults and exceptions as the reference.  What
// differs is underneath: every operator is a gfx950 kernel enqueued on the
// solver's own HIP stream, each cycle is recorded once as a HIP graph and
// replayed (Settings::use_graphs), and the per-cycle residual norm is one
// fused residual+max kernel with a single scalar read-back.
#ifndef POISSON_LIBRARY
#define POISSON_LIBRARY

#include <chrono>
#include <cstring>
#include <tuple>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "boundary.h"
#include "domain.h"
#include "fcycle.h"
#include "fullweighting.h"
#include "gaussseidel.h"
#include "grid.h"
#include "injection.h"
#include "jacobi.h"
#include "trilinearinterpolation.h"
#include "vcycle.h"

namespace Poisson {

template <class T, template <class> class R, template <class> class P, template <class> class S>
class PoissonSolver {
private:
    uint_t num_devices;
    Domain<T>*** domains = nullptr;
    Settings settings;
    BoundaryCondition BC;
    Grid grid;
    R<T> restriction_type;
    P<T> prolongation;
    S<T> relaxation;
    bool is_verbose = false;
    uint_t iter = 0;
    double_t rel_res = 0.0;
    double_t wtime = 0.0;
    int_t dev_shape[3];
    void alloc();

    // ---- graph replay of whole cycles
    void* stream = nullptr;
    bool owns_stream = true;
    detail::Scalar* scalar = nullptr;
    struct LevelState {
        DeviceArray<T>* u;
        DeviceArray<T>* uprev;
        T* devptr;
        T* alt;
        bool zero;
    };
    struct State {
        std::vector<LevelState> levels;
        int relax_state;
        bool operator<(const State& o) const {
            if (relax_state != o.relax_state) return relax_state < o.relax_state;
            for (size_t l = 0; l < levels.size(); l++) {
                const LevelState &a = levels[l], &b = o.levels[l];
                if (a.u != b.u) return a.u < b.u;
                if (a.devptr != b.devptr) return a.devptr < b.devptr;
                if (a.alt != b.alt) return a.alt < b.alt;
                if (a.zero != b.zero) return a.zero < b.zero;
            }
            return false;
        }
    };
    struct Replay {
        void* exec;
        State after;
    };
    // (cycle kind, nsmooth, omega bits, state before) -> graph
    std::map<std::tuple<int, long, uint64_t, State>, Replay> graphs;
    State capture_state() const;
    void apply_state(const State& s);
    void run_cycle(int kind, T omega, int_t nsmooth);
    void cycle_direct(int kind, T omega, int_t nsmooth);
    double level0_residual_norm(bool store = true);
    void clear_graphs();

public:
    PoissonSolver(int_t _dev_shape[3], Settings& _settings, BoundaryCondition& _BC);
    PoissonSolver(Settings& _settings, BoundaryCondition& _BC);
    PoissonSolver(int_t _dev_shape[3], Settings& _settings);
    PoissonSolver(const PoissonSolver&) = delete;
    PoissonSolver& operator=(const PoissonSolver&) = delete;
    ~PoissonSolver();
    void init();
    void init_zero();
    void to_device();
    // to_device() of a freshly constructed (all-zero) solver, made on the device
    void to_device_zero();
    void to_host();
    void verbose(bool onoff);
    void solve(const std::string solver = "Vcycle", const int_t nsmooth = 10, T omega = -1.0, std::string ofile = "",
               double maxtime = 20 * 60);
    void save(uint_t gpuid, std::string file_name = "u.vtk");
    void save_all(uint_t gpuid, std::string ufile_name = "u.vtk", std::string ffile_name = "f.vtk",
                  std::string rfile_name = "r.vtk");
    void save_all(std::string ufile_name = "u.vtk", std::string ffile_name = "f.vtk", std::string rfile_name = "r.vtk");
    T relative_residual();
    T solve_time();
    T solve_iterations();
    DeviceArray<T>& get_rhs(uint_t gpuid);
    const DeviceArray<T>& get_solution(uint_t gpuid);
    const DeviceArray<T>& get_residual(uint_t gpuid);
    Boundary<T>& get_top_bc(uint_t gpuid);
    Boundary<T>& get_bottom_bc(uint_t gpuid);

    // ---- additions
    void init(funptr ufun, funptr ffun, funptr dudxfun, funptr dudyfun, funptr dudzfun);
    Domain<T>& get_domain(uint_t gpuid, uint_t level) { return *domains[gpuid][level]; }
    void* get_stream() const { return stream; }
    bool owns_its_stream() const { return owns_stream; }
    uint_t get_levels() const { return settings.levels; }
    const Settings& get_settings() const { return settings; }
    // one cycle ("Vcycle"/"Fcycle") without norms, enqueued (graph-replayed
    // when enabled); returns without waiting
    void cycle(const std::string solver, const int_t nsmooth = 10, T omega = -1.0);
    // residual on level 0 and ||r||inf / ||f||inf (waits)
    double relres();
    // nsmooth smoothing passes on the finest level (enqueued)
    void smooth_finest(int_t nsmooth, T omega) {
        mgpx::StreamScope scope(stream);
        if (omega < 0) omega = relaxation.default_omega();
        relaxation.relax_level(*domains[0][0], omega, nsmooth);
    }
    // enqueue on `s` (owned by the caller) instead of this solver's own stream;
    // recorded cycle graphs are dropped
    // (wait = false: the caller knows the current stream holds none of this
    // solver's work)
    void use_stream(void* s, bool wait = true) {
        if (s == stream) return;
        if (wait) mgp_stream_sync(stream);
        clear_graphs();
        if (owns_stream) mgp_stream_destroy(stream);
        stream = s;
        owns_stream = false;
    }
    // every level's u back to zero with its boundary data, on the device (waits)
    void reset_solution() {
        mgpx::StreamScope scope(stream);
        for (uint_t l = 0; l < settings.levels; l++) domains[0][l]->reset_solution();
        mgpx::check(mgp_stream_sync(stream), "reset_solution");
    }
};

template <class T, template <class> class R, template <class> class P, template <class> class S>
PoissonSolver<T, R, P, S>::PoissonSolver(int_t _dev_shape[3], Settings& _settings, BoundaryCondition& _BC)
    : num_devices(_dev_shape[0] * _dev_shape[1] * _dev_shape[2]), settings(_settings), BC(_BC),
      grid(_settings, _settings.levels, _settings.print_settings) {
    for (int d = 0; d < 3; d++) dev_shape[d] = _dev_shape[d];
    this->alloc();
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
PoissonSolver<T, R, P, S>::PoissonSolver(Settings& _settings, BoundaryCondition& _BC)
    : num_devices(1), settings(_settings), BC(_BC), grid(_settings, _settings.levels, _settings.print_settings) {
    dev_shape[0] = dev_shape[1] = dev_shape[2] = 1;
    this->alloc();
}

// default BC {N,N,N,N,D,D} (libpoisson.h:89-98)
template <class T, template <class> class R, template <class> class P, template <class> class S>
PoissonSolver<T, R, P, S>::PoissonSolver(int_t _dev_shape[3], Settings& _settings)
    : num_devices(_dev_shape[0] * _dev_shape[1] * _dev_shape[2]), settings(_settings),
      BC({NEUMANN, NEUMANN, NEUMANN, NEUMANN, DIRICHLET, DIRICHLET}), grid(_settings, _settings.levels, _settings.print_settings) {
    for (int d = 0; d < 3; d++) dev_shape[d] = _dev_shape[d];
    this->alloc();
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
PoissonSolver<T, R, P, S>::~PoissonSolver() {
    // queued work first, then the arrays it uses (an adopted stream is the
    // adopter's to drain: it may carry other solvers' work)
    if (stream && owns_stream) mgp_stream_sync(stream);
    clear_graphs();
    if (domains) {
        for (uint_t gpuid = 0; gpuid < num_devices; gpuid++) {
            if (!domains[gpuid]) continue;
            for (uint_t l = 0; l < settings.levels; l++) delete domains[gpuid][l];
            delete[] domains[gpuid];
        }
        delete[] domains;
    }
    delete scalar;
    if (stream && owns_stream) mgp_stream_destroy(stream);
}

// libpoisson.h:113-156 on one device
template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::alloc() {
    if (num_devices != 1)
        throw std::invalid_argument("PoissonSolver: dev_shape must be 1x1x1 (one GPU per solver)");
    mgpx::require_device();
    mgpx::check(mgp_set_device((int)settings.dev), "PoissonSolver: set device");
    if (void* s = mgpx::adopted_stream_slot()) {
        stream = s;
        owns_stream = false;
    } else {
        mgpx::check(mgp_stream_create(&stream), "PoissonSolver: stream");
    }
    mgpx::StreamScope scope(stream);
    scalar = new detail::Scalar();
    domains = new Domain<T>**[1];
    domains[0] = new Domain<T>*[settings.levels]();
    int_t device[3] = {0, 0, 0};
    for (uint_t l = 0; l < settings.levels; l++) {
        grid.domainsettings[l].dev = settings.dev;
        domains[0][l] = new Domain<T>(grid.domainsettings[l], BC, relaxation.requires_duplicate_solution(), device,
                                      dev_shape);
    }
    domains[0][settings.levels - 1]->coarsest = true;
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::init() {
    this->init(&ufun, &ffun, &dudxfun, &dudyfun, &dudzfun);
}

// libpoisson.h:159-171: analytic data on level 0, zeros elsewhere
template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::init(funptr u, funptr f, funptr dx, funptr dy, funptr dz) {
    mgpx::StreamScope scope(stream);
    for (uint_t l = 0; l < settings.levels; l++) {
        if (l == 0)
            domains[0][l]->init(u, f, dx, dy, dz);
        else
            domains[0][l]->init_zero();
    }
    mgpx::check(mgp_stream_sync(stream), "init");
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::init_zero() {
    mgpx::StreamScope scope(stream);
    for (uint_t l = 0; l < settings.levels; l++) domains[0][l]->init_zero();
    mgpx::check(mgp_stream_sync(stream), "init_zero");
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::to_device_zero() {
    mgpx::StreamScope scope(stream);
    clear_graphs();
    for (uint_t l = 0; l < settings.levels; l++) domains[0][l]->to_device_zero();
    for (uint_t l = 1; l < settings.levels; l++) domains[0][l]->prepare_coarse_faces();
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::to_device() {
    mgpx::StreamScope scope(stream);
    clear_graphs();
    for (uint_t l = 0; l < settings.levels; l++) domains[0][l]->to_device();
    for (uint_t l = 1; l < settings.levels; l++) domains[0][l]->prepare_coarse_faces();
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::to_host() {
    mgpx::StreamScope scope(stream);
    for (uint_t l = 0; l < settings.levels; l++) domains[0][l]->to_host();
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::verbose(bool onoff) {
    this->is_verbose = onoff;
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
typename PoissonSolver<T, R, P, S>::State PoissonSolver<T, R, P, S>::capture_state() const {
    State s;
    s.relax_state = relaxation.mgp_state();
    for (uint_t l = 0; l < settings.levels; l++) {
        Domain<T>& d = *domains[0][l];
        s.levels.push_back(LevelState{d.u, d.uprev, d.u->devptr, d.u_alt, d.u_zero});
    }
    return s;
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::apply_state(const State& s) {
    relaxation.mgp_set_state(s.relax_state);
    for (uint_t l = 0; l < settings.levels; l++) {
        Domain<T>& d = *domains[0][l];
        d.u = s.levels[l].u;
        d.uprev = s.levels[l].uprev;
        d.u->devptr = s.levels[l].devptr;
        d.u_alt = s.levels[l].alt;
        d.u_zero = s.levels[l].zero;
    }
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::clear_graphs() {
    for (auto& kv : graphs) mgp_graph_destroy(kv.second.exec);
    graphs.clear();
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::cycle_direct(int kind, T omega, int_t nsmooth) {
    if (kind == 1)
        Vcycle<T>(this->domains, this->restriction_type, this->prolongation, this->relaxation, omega, 0,
                  settings.levels, num_devices, nsmooth);
    else if (kind == 2)
        Fcycle<T>(this->domains, this->restriction_type, this->prolongation, this->relaxation, omega, 0,
                  settings.levels, num_devices, nsmooth);
}

// One cycle: replayed from a graph recorded for the current host-side state
// (which buffer is u on each level), or run directly.  Recording a cycle
// advances the host state exactly as running it does; a replay then applies
// the state recorded after it.
template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::run_cycle(int kind, T omega, int_t nsmooth) {
    if (kind == 0) return;  // "relaxation": no work per iteration (libpoisson.h:304-308)
    if (!settings.use_graphs) {
        cycle_direct(kind, omega, nsmooth);
        return;
    }
    // omega and nsmooth are baked into a graph: key on them too
    const State before = capture_state();
    const double od = (double)omega;
    uint64_t obits = 0;
    std::memcpy(&obits, &od, sizeof(obits));
    const auto key = std::make_tuple(kind, (long)nsmooth, obits, before);
    auto it = graphs.find(key);
    if (it != graphs.end()) {
        mgpx::check(mgp_graph_launch(it->second.exec, stream), "graph launch");
        apply_state(it->second.after);
        return;
    }
    void* exec = nullptr;
    mgpx::check(mgp_graph_begin(stream), "graph capture");
    try {
        cycle_direct(kind, omega, nsmooth);
    } catch (...) {
        void* junk = nullptr;
        mgp_graph_end(stream, &junk);
        mgp_graph_destroy(junk);
        throw;
    }
    mgpx::check(mgp_graph_end(stream, &exec), "graph capture");
    graphs[key] = Replay{exec, capture_state()};
    // the recording enqueued nothing; run it once now
    const State after = capture_state();
    apply_state(before);
    mgpx::check(mgp_graph_launch(exec, stream), "graph launch");
    apply_state(after);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
double PoissonSolver<T, R, P, S>::level0_residual_norm(bool store) {
    mgpx::check(mgp_zero_scalars(scalar->dev, 1, stream), "residual norm");
    if (store) residual_norm<T>(*domains[0][0], scalar->dev);
    else residual_norm_only<T>(*domains[0][0], scalar->dev);
    mgpx::check(mgp_memcpy_d2h(scalar->host, scalar->dev, sizeof(double), stream), "residual norm");
    mgpx::check(mgp_stream_sync(stream), "residual norm");
    return scalar->host[0];
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::cycle(const std::string solver, const int_t nsmooth, T omega) {
    mgpx::StreamScope scope(stream);
    const bool v = solver == "vcycle" || solver == "Vcycle";
    const bool f = solver == "fcycle" || solver == "Fcycle";
    const bool r = solver == "relaxation" || solver == "Relaxation";
    if (!(v || f || r))
        throw std::invalid_argument("solver must be \"Relaxation\", \"Vcycle\" of \"Fcycle\" but was \"" + solver + "\"");
    if (f && settings.levels < 2) throw std::invalid_argument("Fcycle needs at least 2 levels");
    if (omega < 0) omega = relaxation.default_omega();
    run_cycle(v ? 1 : (f ? 2 : 0), omega, nsmooth);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
double PoissonSolver<T, R, P, S>::relres() {
    mgpx::StreamScope scope(stream);
    const double fnorm = domains[0][0]->f->infinity_norm();
    return level0_residual_norm() / fnorm;
}

// libpoisson.h:237-357
template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::solve(const std::string solver, const int_t nsmooth, T omega, std::string ofile,
                                      double maxtime) {
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto since = [](clock::time_point a) { return std::chrono::duration<double>(clock::now() - a).count(); };
    mgpx::StreamScope scope(stream);
    const bool use_vcycle = solver == "vcycle" || solver == "Vcycle";
    const bool use_fcycle = solver == "fcycle" || solver == "Fcycle";
    const bool use_relaxation = solver == "relaxation" || solver == "Relaxation";
    if (!(use_vcycle || use_relaxation || use_fcycle))
        throw std::invalid_argument("solver must be \"Relaxation\", \"Vcycle\" of \"Fcycle\" but was \"" + solver + "\"");
    if (use_fcycle && settings.levels < 2) throw std::invalid_argument("Fcycle needs at least 2 levels");
    if (omega < 0) omega = relaxation.default_omega();
    const int kind = use_vcycle ? 1 : (use_fcycle ? 2 : 0);

    iter = 0;
    const double_t fnorm = domains[0][0]->f->infinity_norm();
    // the initial residual refreshes no neighbour data (libpoisson.h:268-271)
    mgpx::check(mgp_zero_scalars(scalar->dev, 1, stream), "residual norm");
    residual_norm<T>(*domains[0][0], scalar->dev, false);
    mgpx::check(mgp_memcpy_d2h(scalar->host, scalar->dev, sizeof(double), stream), "residual norm");
    mgpx::check(mgp_stream_sync(stream), "residual norm");
    const double_t rnorm = scalar->host[0];
    rel_res = rnorm / fnorm;

    if (is_verbose) std::cout << std::setw(4) << 0 << ": Initial residual: " << std::setw(8) << rel_res << std::endl;
    std::ofstream out;
    if (ofile.compare("") != 0) {
        out.open(ofile, std::ios::app);
        out << "#    seconds     rel_res" << std::endl;
    }
    const bool consistent = settings.mode == CONSISTENT;
    for (iter = 0; iter < settings.maxiter; iter++) {
        const auto tc = clock::now();
        run_cycle(kind, omega, nsmooth);
        // the norm only: r is written once, below, for the final u
        T norm = (T)level0_residual_norm(false);
        if (settings.print_cycle_time && kind != 0)
            std::cout << (kind == 1 ? "Vcycle" : "Fcycle") << " took " << since(tc) << " seconds" << std::endl;
        norm /= fnorm;
        if (ofile.compare("") != 0) {
            out << std::setw(12) << since(t0);
            out << std::setw(12) << this->relative_residual();
            out << std::endl;
        }
        if (is_verbose) std::cout << std::setw(4) << iter + 1 << ": Relative residual: " << std::setw(8) << norm << std::endl;
        if (consistent && norm <= settings.tolerance) {
            rel_res = norm;
            break;
        }
        if (iter >= settings.miniter) {  // libpoisson.h:340-353
            if (norm > 2.0 * rel_res) {
                rel_res = norm;
                break;
            } else if (!consistent && std::abs(norm - rel_res) < this->settings.tolerance) {
                rel_res = norm;
                break;
            }
            if (since(t0) > maxtime) {
                std::cout << "WARNING: Solver reached maximum wall time without converging!" << std::endl;
                break;
            }
        }
        rel_res = norm;
    }
    if (settings.maxiter > 0) {
        // r of the final u (what the last iteration's residual left in the reference)
        residual<T>(*domains[0][0]);
        mgpx::check(mgp_stream_sync(stream), "residual");
    }
    wtime = since(t0);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::save(uint_t gpuid, std::string file_name) {
    domains[gpuid][0]->save(file_name);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::save_all(uint_t gpuid, std::string u, std::string f, std::string r) {
    domains[gpuid][0]->save_all(u, f, r);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
void PoissonSolver<T, R, P, S>::save_all(std::string u, std::string f, std::string r) {
    domains[0][0]->save_all(u, f, r);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
T PoissonSolver<T, R, P, S>::relative_residual() {
    return rel_res;
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
T PoissonSolver<T, R, P, S>::solve_time() {
    return wtime;
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
T PoissonSolver<T, R, P, S>::solve_iterations() {
    return iter;
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
DeviceArray<T>& PoissonSolver<T, R, P, S>::get_rhs(uint_t gpuid) {
    return *(domains[gpuid][0]->f);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
const DeviceArray<T>& PoissonSolver<T, R, P, S>::get_solution(uint_t gpuid) {
    return *(domains[gpuid][0]->u);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
const DeviceArray<T>& PoissonSolver<T, R, P, S>::get_residual(uint_t gpuid) {
    return *(domains[gpuid][0]->r);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
Boundary<T>& PoissonSolver<T, R, P, S>::get_top_bc(uint_t gpuid) {
    return *(domains[gpuid][0]->top);
}

template <class T, template <class> class R, template <class> class P, template <class> class S>
Boundary<T>& PoissonSolver<T, R, P, S>::get_bottom_bc(uint_t gpuid) {
    return *(domains[gpuid][0]->bottom);
}
}  // namespace Poisson

// the reference's headers leave these names in the global namespace
// (libpoisson.h:19-25, array.h:78-82); sources written against it use them
#include "refinement.h"

using std::cerr;
using std::cout;
using std::endl;
using std::ios;
using std::ofstream;
using std::setw;
using std::string;
using Poisson::DIRICHLET;
using Poisson::NEUMANN;
#endif
