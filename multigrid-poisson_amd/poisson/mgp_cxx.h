// C++ glue over the mgp C ABI: error checking, the current stream, and
// dtype-dispatching wrappers used by the Poisson classes.
#ifndef POISSON_MGP_CXX
#define POISSON_MGP_CXX

#include <stdexcept>
#include <string>

#include "arena.h"
#include "definitions.h"
#include "mgp.h"

namespace Poisson {
namespace mgpx {

inline void check(int rc, const char* what) {
    if (rc != MGP_OK) {
        if (rc == MGP_ERR_ARG) throw std::invalid_argument(std::string(what) + ": " + mgp_error_string(rc));
        throw std::runtime_error(std::string(what) + ": " + mgp_error_string(rc));
    }
}

// The stream the Poisson classes enqueue on.  A solver installs its own with
// StreamScope for the duration of each call; otherwise a process-wide library
// stream is used.
inline void*& current_stream_slot() {
    thread_local void* s = nullptr;
    return s;
}

inline void* default_stream() {
    static void* s = [] {
        int n = 0;
        if (mgp_device_count(&n) != MGP_OK || n < 1)
            throw std::runtime_error("Device array requested but no GPU device was detected\n");
        void* st = nullptr;
        check(mgp_stream_create(&st), "mgp_stream_create");
        return st;
    }();
    return s;
}

inline void* stream() {
    void* s = current_stream_slot();
    return s ? s : default_stream();
}

struct StreamScope {
    void* prev;
    explicit StreamScope(void* s) : prev(current_stream_slot()) { current_stream_slot() = s; }
    ~StreamScope() { current_stream_slot() = prev; }
};

// A stream for the solvers constructed on this thread to enqueue on instead
// of creating their own (owned by whoever installs it, e.g. a server worker
// that runs many solvers in turn); nullptr: each solver creates its own.  A
// solver on an adopted stream does not wait for it when destroyed: the
// adopter makes sure the solver's queued work has finished first.
inline void*& adopted_stream_slot() {
    thread_local void* s = nullptr;
    return s;
}

struct AdoptStream {
    void* prev;
    explicit AdoptStream(void* s) : prev(adopted_stream_slot()) { adopted_stream_slot() = s; }
    ~AdoptStream() { adopted_stream_slot() = prev; }
    AdoptStream(const AdoptStream&) = delete;
    AdoptStream& operator=(const AdoptStream&) = delete;
};

inline void require_device() {
    int n = 0;
    if (mgp_device_count(&n) != MGP_OK || n < 1)
        throw std::runtime_error("Device array requested but no GPU device was detected\n");
}

// dtype dispatch
template <class T> struct Ops;

#define MGPX_OPS(T, SFX)                                                                                          \
    template <> struct Ops<T> {                                                                                   \
        static int fill(const mgp_array* a, double v, void* s) { return mgp_fill_##SFX(a, v, s); }               \
        static int add(const mgp_array* a, const mgp_array* b, void* s) { return mgp_add_##SFX(a, b, s); }       \
        static int norm_inf(const mgp_array* a, double* o, void* s) { return mgp_norm_inf_##SFX(a, o, s); }      \
        static int rbgs_half(int m, const mgp_array* u, const mgp_array* f, const int64_t b[6], double w,        \
                             double h, int c, void* s) {                                                         \
            return mgp_rbgs_half_##SFX(m, u, f, b, w, h, c, s);                                                   \
        }                                                                                                         \
        static int rbgs_sweep(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f,               \
                              const mgp_faces* F, const int64_t b[6], double w, void* s) {                       \
            return mgp_rbgs_sweep_##SFX(m, ui, uo, f, F, b, w, s);                                                \
        }                                                                                                         \
        static int residual_norm(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r,              \
                                 const int64_t b[6], double h, double* n, void* s) {                            \
            return mgp_residual_norm_##SFX(m, u, f, r, b, h, n, s);                                               \
        }                                                                                                         \
        static int residual_restrict_fw(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r,       \
                                        const mgp_array* c, const int64_t b[6], double h, void* s) {             \
            return mgp_residual_restrict_fw_##SFX(m, u, f, r, c, b, h, s);                                        \
        }                                                                                                         \
        static int rbgs_sweep2(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f,              \
                               const mgp_faces* F, const int64_t b[6], double w, void* s) {                      \
            return mgp_rbgs_sweep2_##SFX(m, ui, uo, f, F, b, w, s);                                               \
        }                                                                                                         \
        static int rbgs_sweep2_prolong(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f,      \
                                       const mgp_faces* F, const int64_t b[6], double w, const mgp_array* c,     \
                                       void* s) {                                                                \
            return mgp_rbgs_sweep2_prolong_##SFX(m, ui, uo, f, F, b, w, c, s);                                   \
        }                                                                                                        \
        static int rbgs_sweep2_from_zero(int m, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,     \
                                         const int64_t b[6], double w, void* s) {                                \
            return mgp_rbgs_sweep2_from_zero_##SFX(m, uo, f, F, b, w, s);                                        \
        }                                                                                                        \
        static int rbgs_sweeps(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f,              \
                               const mgp_faces* F, const int64_t b[6], double w, int64_t n, void* s) {           \
            return mgp_rbgs_sweeps_##SFX(m, ui, uo, f, F, b, w, n, s);                                            \
        }                                                                                                         \
        static int jacobi(int m, const mgp_array* u, const mgp_array* v, const mgp_array* f, const int64_t b[6], \
                          double w, double h, void* s) {                                                         \
            return mgp_jacobi_##SFX(m, u, v, f, b, w, h, s);                                                      \
        }                                                                                                         \
        static int residual(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r,                   \
                            const int64_t b[6], double h, double* n, void* s) {                                  \
            return mgp_residual_##SFX(m, u, f, r, b, h, n, s);                                                    \
        }                                                                                                         \
        static int inject(const mgp_array* i, const mgp_array* o, void* s) {                                     \
            return mgp_restrict_inject_##SFX(i, o, s);                                                            \
        }                                                                                                         \
        static int fw(const mgp_array* i, const mgp_array* o, void* s) { return mgp_restrict_fw_##SFX(i, o, s); } \
        static int prolong(const mgp_array* c, const mgp_array* f, void* s) {                                    \
            return mgp_prolong_trilinear_##SFX(c, f, s);                                                          \
        }                                                                                                         \
        static int prolong_add(const mgp_array* c, const mgp_array* u, void* s) {                                \
            return mgp_prolong_add_##SFX(c, u, s);                                                                \
        }                                                                                                         \
        static int bc_dirichlet(const mgp_array* u, int l, const void* g, void* s) {                             \
            return mgp_bc_dirichlet_##SFX(u, l, g, s);                                                            \
        }                                                                                                         \
        static int bc_neumann(const mgp_array* u, int l, const void* g, double h, void* s) {                     \
            return mgp_bc_neumann_##SFX(u, l, g, h, s);                                                           \
        }                                                                                                         \
        static int bc_neumann_faces(const mgp_array* u, const mgp_faces* F, void* s) {                           \
            return mgp_bc_neumann_faces_##SFX(u, F, s);                                                           \
        }                                                                                                         \
        static int bc_restrict_neumann(int r, const mgp_array* u, int l, const void* gf, void* gc, double h,     \
                                       void* s) {                                                                \
            return mgp_bc_restrict_neumann_##SFX(r, u, l, gf, gc, h, s);                                          \
        }                                                                                                         \
    };

MGPX_OPS(double, f64)
MGPX_OPS(float, f32)
#undef MGPX_OPS

}  // namespace mgpx
}  // namespace Poisson
#endif
