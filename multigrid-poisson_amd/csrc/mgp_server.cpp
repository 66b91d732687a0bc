// Serving loop (mgp_server_*): Poisson solves of mixed shapes and boundary
// types accepted as they arrive and run concurrently on one GPU.  No
// counterpart in the reference, which runs one solve per process.
//
// max_active worker threads, each with a stream of its own.  A worker takes
// the oldest pending request once the HBM budget allows its hierarchy (a
// RefinementSolver: FP32 F-cycles inside FP64 iterative refinement, on the
// worker's stream), solves it with the same solve() a lone caller would run,
// copies the solution to the caller's buffer and takes the next request.
// Hierarchies stay resident after their request ends and are reused by later
// requests of the same shape and face types; idle ones are evicted least
// recently used first when a new shape needs room.  Requests that arrive
// while others run are admitted as workers and memory free up; their kernels
// overlap on the device (one stream each) and their host work (hierarchy
// set-up, graph recording, result copies) overlaps across worker threads.  A
// request's arithmetic does not depend on what else runs (own arrays and
// stream, deterministic kernels), so its result is the same bits alone or in
// a mix.
//
// Hierarchies are carved from one device region allocated when the server
// starts (and their pinned scalars from one pinned host region), so building
// one for a new shape and evicting an idle one make no hipMalloc / hipFree:
// hipFree waits for the whole device, and with a mix of shapes evictions
// were frequent enough to stall every running solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "libpoisson.h"
#include "mgp.h"

namespace mgp {
void set_last_error(const char* msg);
}

namespace {

using Clock = std::chrono::steady_clock;
using Key = std::tuple<int64_t, int64_t, int64_t, uint32_t>;  // dims, face-type bits

double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Job {
    int64_t id = 0;
    mgp_request req{};
    Clock::time_point t_submit, t_admit, t_done;
    mgp_result res{};
    bool done = false;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// levels down to a 3-point coarsest axis with exact 2:1 nesting
int64_t max_levels(const int64_t d[3]) {
    int64_t L = 1;
    while (true) {
        bool ok = true;
        for (int a = 0; a < 3; a++) {
            const int64_t s = (d[a] - 1) >> (L - 1);
            if (s % 2 != 0 || s / 2 < 2) ok = false;
        }
        if (!ok) return L;
        L++;
    }
}

struct SlotBase {
    Key key;
    size_t bytes = 0;
    Job* job = nullptr;
    virtual ~SlotBase() {}
    virtual void begin(const mgp_request& r, void* stream) = 0;  // on `stream`: f, u = 0 (waits)
    virtual void run() = 0;                         // the solve (waits)
    virtual void finish(double* out) = 0;           // solution to the host (waits)
    virtual void stats(int64_t* iters, double* rel) = 0;
};

template <template <class> class S>
struct Slot : SlotBase {
    Poisson::Settings st;
    Poisson::BoundaryCondition bc;
    std::unique_ptr<Poisson::RefinementSolver<Poisson::FullWeighting, Poisson::TrilinearInterpolation, S>> solver;
    int64_t nsmooth;
    double omega;

    Slot(const mgp_request& r, const mgp_server_config& c) {
        for (int d = 0; d < 3; d++) {
            st.dims[d] = r.dims[d];
            st.origin[d] = 0.0;
        }
        st.levels = max_levels(r.dims);
        st.h = 1.0 / (double)(r.dims[0] - 1);
        st.lengthx = 1.0;
        st.maxiter = c.maxiter;
        st.miniter = 1;
        st.tolerance = c.tolerance;
        st.coarse_sweeps = c.coarse_sweeps;
        st.mode = Poisson::CONSISTENT;
        // the coarsest level solved by CG (mgp_cg_solve_*): with 2^k+1 axes
        // drawn independently the coarsest level can keep a 33-point axis,
        // whose smooth error smoothing passes barely reduce
        st.coarse_cg = true;
        // a new hierarchy (a shape not seen recently) serves a few iterations:
        // recording and instantiating its cycle graphs costs more than the
        // launches they save, unless the caller asks for graphs
        st.use_graphs = (c.flags & MGP_SERVER_GRAPHS) != 0;
        st.print_cycle_time = false;
        st.print_settings = false;
        auto t = [&](int q) { return r.bc[q] == MGP_NEUMANN ? Poisson::NEUMANN : Poisson::DIRICHLET; };
        bc = Poisson::BoundaryCondition{t(MGP_NORTH), t(MGP_SOUTH), t(MGP_EAST), t(MGP_WEST), t(MGP_TOP), t(MGP_BOTTOM)};
        solver.reset(new Poisson::RefinementSolver<Poisson::FullWeighting, Poisson::TrilinearInterpolation, S>(st, bc));
        solver->to_device_zero();  // zero f, u and (zero) face data, on the device
        nsmooth = c.nsmooth;
        omega = c.omega;
    }
    void begin(const mgp_request& r, void* stream) override {
        // an idle hierarchy has no queued work: it moves to this worker's
        // stream without waiting for the stream it last ran on
        solver->use_stream(stream, false);
        Poisson::Domain<double>& d = solver->fp64().get_domain(0, 0);
        const mgp_array f = d.f->desc();
        Poisson::mgpx::check(mgp_fill_modes_f64(&f, r.nmodes, r.abc, r.amp, st.h, solver->get_stream()), "rhs");
        solver->reset_solution();  // u = 0 with the (zero) boundary data; waits
    }
    void run() override { solver->solve("Fcycle", nsmooth, omega, "", 1e30); }
    void finish(double* out) override {
        if (!out) return;
        // pack the interior into the (now free) FP64 residual array, then one copy
        Poisson::Domain<double>& d = solver->fp64().get_domain(0, 0);
        const mgp_array u = d.u->desc();
        void* st = solver->get_stream();
        Poisson::mgpx::check(mgp_pack_f64(&u, d.r->devptr, st), "server result");
        const size_t n = (size_t)(u.shape[0] * u.shape[1] * u.shape[2]);
        hip_check(hipMemcpyAsync(out, d.r->devptr, n * sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)st),
                  "server result");
        hip_check(hipStreamSynchronize((hipStream_t)st), "server result");
    }
    void stats(int64_t* iters, double* rel) override {
        *iters = (int64_t)solver->solve_iterations();
        *rel = solver->relative_residual();
    }
};

class Server {
public:
    explicit Server(const mgp_server_config& c) : cfg(c) {
        if (c.hbm_budget <= 0 || c.max_active < 1 || c.nsmooth < 1 || c.maxiter < 1 || !(c.tolerance >= 0.0) ||
            (c.relaxation != MGP_GAUSS_SEIDEL && c.relaxation != MGP_JACOBI) || !std::isfinite(c.omega))
            throw std::invalid_argument("invalid server config");
        // omega <= 0 stands for the smoother's default (include/jacobi.h:29-31,
        // include/gaussseidel.h:97-99 of the reference); the solver itself only
        // treats negative values so, and omega = 0 would smooth nothing
        if (cfg.omega <= 0.0) cfg.omega = c.relaxation == MGP_JACOBI ? 6.0 / 7.0 : 8.0 / 9.0;
        int dev = 0;
        hip_check(hipGetDevice(&dev), "server");
        if (!(c.flags & MGP_SERVER_DIRECT_ALLOC)) make_arenas();
        streams.resize((size_t)c.max_active, nullptr);
        try {
            for (auto& st : streams) hip_check(hipStreamCreateWithFlags((hipStream_t*)&st, hipStreamNonBlocking),
                                               "server stream");
        } catch (...) {
            release();
            throw;
        }
        for (int w = 0; w < c.max_active; w++) workers.emplace_back([this, dev, w] { run(dev, streams[(size_t)w]); });
    }
    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stopping = true;
        }
        cv_work.notify_all();
        for (auto& t : workers) t.join();
        idle.clear();  // every slot's work has finished (workers wait for it)
        release();
    }

    int64_t submit(const mgp_request& r) {
        for (int d = 0; d < 3; d++)
            if (r.dims[d] < 3) throw std::invalid_argument("dims must be >= 3");
        bool any_dirichlet = false;
        for (int q = 0; q < 6; q++) {
            if (r.bc[q] != MGP_DIRICHLET && r.bc[q] != MGP_NEUMANN) throw std::invalid_argument("bad face type");
            any_dirichlet |= r.bc[q] == MGP_DIRICHLET;
        }
        // all-Neumann problems are singular (solution up to a constant)
        if (!any_dirichlet) throw std::invalid_argument("all-Neumann requests are not accepted");
        if (r.nmodes < 0 || r.nmodes > MGP_MAX_MODES) throw std::invalid_argument("nmodes out of range");
        int64_t tmp[3] = {r.dims[0], r.dims[1], r.dims[2]};
        if (max_levels(tmp) < 2) throw std::invalid_argument("shape does not coarsen (need 2^k m + 1 axes)");
        auto j = std::make_unique<Job>();
        j->req = r;
        j->t_submit = Clock::now();
        std::lock_guard<std::mutex> lk(mu);
        j->id = next_id++;
        const int64_t id = j->id;
        pending.push_back(j.get());
        jobs[id] = std::move(j);
        cv_work.notify_all();
        return id;
    }

    mgp_result wait(int64_t id) {
        std::unique_lock<std::mutex> lk(mu);
        auto it = jobs.find(id);
        if (it == jobs.end()) throw std::invalid_argument("unknown request id");
        Job* j = it->second.get();
        cv_done.wait(lk, [&] { return j->done; });
        const mgp_result r = j->res;
        jobs.erase(it);
        return r;
    }

    void stats(int64_t* completed_out, int64_t* peak_out, int64_t* resident_out) {
        std::lock_guard<std::mutex> lk(mu);
        if (completed_out) *completed_out = completed;
        if (peak_out) *peak_out = peak_active;
        if (resident_out) *resident_out = (int64_t)bytes_used;
    }

private:
    mgp_server_config cfg;
    std::vector<void*> streams;  // one per worker, destroyed after every slot
    void* dev_region = nullptr;
    void* host_region = nullptr;
    std::unique_ptr<Poisson::mgpx::Arena> dev_arena, host_arena;
    std::vector<std::thread> workers;
    std::mutex mu;  // guards everything below
    std::condition_variable cv_work, cv_done;
    bool stopping = false;
    int64_t next_id = 1;
    std::deque<Job*> pending;
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    std::list<std::unique_ptr<SlotBase>> idle;  // resident, least recently used first
    size_t bytes_used = 0;                      // active + resident hierarchies (and reservations)
    int64_t active = 0, completed = 0, peak_active = 0;

    // the device region: the budget plus an eighth for the arena's
    // fragmentation, within what the device has free (halved until it can be
    // allocated; without it hierarchies are allocated one array at a time)
    void make_arenas() {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return;
        const size_t margin = (size_t)1 << 30;
        size_t want = (size_t)cfg.hbm_budget + (size_t)cfg.hbm_budget / 8;
        if (free_b > margin) want = std::min(want, free_b - margin);
        for (; want >= ((size_t)256 << 20); want /= 2)
            if (hipMalloc(&dev_region, want) == hipSuccess) break;
        (void)hipGetLastError();
        if (!dev_region) return;
        dev_arena.reset(new Poisson::mgpx::Arena(dev_region, want));
        Poisson::mgpx::register_arena(dev_arena.get());
        const size_t host_bytes = (size_t)4 << 20;  // 16384 pinned scalars
        if (hipHostMalloc(&host_region, host_bytes, hipHostMallocDefault) == hipSuccess) {
            host_arena.reset(new Poisson::mgpx::Arena(host_region, host_bytes));
            Poisson::mgpx::register_arena(host_arena.get());
        } else {
            host_region = nullptr;
            (void)hipGetLastError();
        }
    }
    void release() {
        for (auto& st : streams)
            if (st) (void)hipStreamDestroy((hipStream_t)st);
        streams.clear();
        if (dev_arena) Poisson::mgpx::unregister_arena(dev_arena.get());
        if (host_arena) Poisson::mgpx::unregister_arena(host_arena.get());
        dev_arena.reset();
        host_arena.reset();
        if (dev_region) (void)hipFree(dev_region);
        if (host_region) (void)hipHostFree(host_region);
        dev_region = host_region = nullptr;
    }

    static Key key_of(const mgp_request& r) {
        uint32_t bits = 0;
        for (int q = 0; q < 6; q++) bits |= (uint32_t)(r.bc[q] == MGP_NEUMANN) << q;
        return Key{r.dims[0], r.dims[1], r.dims[2], bits};
    }
    // device bytes of a request's hierarchy: level 0 in FP64 (u, u_alt, f, r),
    // the two FP32 lo words of its hi + lo iterate and the FP32 hierarchy (u,
    // u_alt, f, r; 8/7 of level 0), halos included
    static size_t need_of(const mgp_request& r) {
        const double pts = (double)(r.dims[0] + 2) * (double)(r.dims[1] + 2) * (double)(r.dims[2] + 2);
        return (size_t)(pts * (4 * 8 + 2 * 4 + 4 * 4 * 8.0 / 7.0) * 1.05) + (4u << 20);
    }

    // under mu: the oldest request and a resident slot for it, or a
    // reservation of its bytes for a new slot; false when nothing can start.
    // Idle slots evicted for room go to `evicted`, destroyed outside the lock.
    bool take(Job*& j, std::unique_ptr<SlotBase>& s, size_t& reserved,
              std::vector<std::unique_ptr<SlotBase>>& evicted) {
        if (pending.empty()) return false;
        j = pending.front();
        const Key k = key_of(j->req);
        for (auto it = idle.begin(); it != idle.end(); ++it)
            if ((*it)->key == k) {
                s = std::move(*it);
                idle.erase(it);
                pending.pop_front();
                reserved = 0;
                return true;
            }
        const size_t need = need_of(j->req);
        while (bytes_used + need > (size_t)cfg.hbm_budget && !idle.empty()) {
            bytes_used -= idle.front()->bytes;
            evicted.push_back(std::move(idle.front()));
            idle.pop_front();
        }
        // wait for room unless nothing else holds memory (then try anyway)
        if (bytes_used + need > (size_t)cfg.hbm_budget && active > 0) return false;
        bytes_used += need;
        reserved = need;
        pending.pop_front();
        return true;
    }

    void finish_job(Job* j, int status) {
        j->t_done = Clock::now();
        std::lock_guard<std::mutex> lk(mu);
        j->res.status = status;
        j->res.latency_s = seconds(j->t_submit, j->t_done);
        j->res.queue_s = seconds(j->t_submit, j->t_admit);
        j->res.run_s = seconds(j->t_admit, j->t_done);
        j->done = true;
        completed++;
        cv_done.notify_all();
    }

    void run(int dev, void* stream) {
        (void)hipSetDevice(dev);
        // this worker's hierarchies: carved from the server's regions and
        // enqueued on its stream
        Poisson::mgpx::ArenaScope arenas(dev_arena.get(), host_arena.get());
        Poisson::mgpx::AdoptStream adopt(stream);
        while (true) {
            Job* j = nullptr;
            std::unique_ptr<SlotBase> s;
            size_t reserved = 0;
            std::vector<std::unique_ptr<SlotBase>> evicted;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return (stopping && pending.empty()) || take(j, s, reserved, evicted); });
                if (!j) return;  // stopping, nothing pending
                active++;
                peak_active = std::max(peak_active, active);
            }
            evicted.clear();  // idle: no queued work, blocks back to the arena
            j->t_admit = Clock::now();
            int status = MGP_OK;
            try {
                if (!s) {
                    s = make_slot(j->req);
                    std::lock_guard<std::mutex> lk(mu);
                    // the reservation now belongs to the slot (s->bytes is the
                    // same estimate, need_of) and leaves with it
                    reserved = 0;
                }
                s->begin(j->req, stream);
                j->res.setup_s = seconds(j->t_admit, Clock::now());
                s->run();
                const auto t0 = Clock::now();
                s->finish(j->req.u_out);
                j->res.fetch_s = seconds(t0, Clock::now());
                s->stats(&j->res.iterations, &j->res.rel_res);
            } catch (const std::exception& e) {
                mgp::set_last_error(e.what());
                // the caller sees only the status: the reason goes to stderr
                std::fprintf(stderr, "mgp server: request %lld (%lldx%lldx%lld) failed: %s\n", (long long)j->id,
                             (long long)j->req.dims[0], (long long)j->req.dims[1], (long long)j->req.dims[2],
                             e.what());
                status = MGP_ERR_STATE;
            }
            // nothing of this request stays queued (a failed one may have
            // left work behind): its slot's blocks can go back to the arena
            (void)hipStreamSynchronize((hipStream_t)stream);
            (void)hipGetLastError();
            // a slot whose request failed may hold a broken graph or state:
            // it is destroyed, not kept for later requests
            std::unique_ptr<SlotBase> broken;
            {
                std::lock_guard<std::mutex> lk(mu);
                bytes_used -= reserved;
                if (s && status == MGP_OK) {
                    idle.push_back(std::move(s));
                } else if (s) {
                    bytes_used -= s->bytes;
                    broken = std::move(s);
                }
                active--;
            }
            broken.reset();
            cv_work.notify_all();  // memory or a worker came free
            finish_job(j, status);
        }
    }

    std::unique_ptr<SlotBase> make_slot(const mgp_request& r) {
        std::unique_ptr<SlotBase> s;
        if (cfg.relaxation == MGP_GAUSS_SEIDEL) s.reset(new Slot<Poisson::GaussSeidel>(r, cfg));
        else s.reset(new Slot<Poisson::Jacobi>(r, cfg));
        s->key = key_of(r);
        s->bytes = need_of(r);
        return s;
    }
};

template <class F>
int server_guarded(F&& f) {
    try {
        f();
        return MGP_OK;
    } catch (const std::invalid_argument& e) {
        mgp::set_last_error(e.what());
        return MGP_ERR_ARG;
    } catch (const std::exception& e) {
        mgp::set_last_error(e.what());
        return MGP_ERR_STATE;
    }
}

}  // namespace

extern "C" {

int mgp_server_create(const mgp_server_config* c, void** out) {
    if (!c || !out) return MGP_ERR_ARG;
    return server_guarded([&] { *out = new Server(*c); });
}
int mgp_server_destroy(void* s) {
    return server_guarded([&] { delete static_cast<Server*>(s); });
}
int mgp_server_submit(void* s, const mgp_request* r, int64_t* id) {
    if (!s || !r || !id) return MGP_ERR_ARG;
    return server_guarded([&] { *id = static_cast<Server*>(s)->submit(*r); });
}
int mgp_server_wait(void* s, int64_t id, mgp_result* out) {
    if (!s || !out) return MGP_ERR_ARG;
    return server_guarded([&] { *out = static_cast<Server*>(s)->wait(id); });
}
int mgp_server_stats(void* s, int64_t* completed, int64_t* peak_active, int64_t* bytes_resident) {
    if (!s) return MGP_ERR_ARG;
    return server_guarded([&] { static_cast<Server*>(s)->stats(completed, peak_active, bytes_resident); });
}

}  // extern "C"
