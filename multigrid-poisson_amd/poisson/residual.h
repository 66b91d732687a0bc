// Residual r = f - L_h u over the relaxation box after the boundary refresh
// (reference include/residual.h:18-81).
#ifndef POISSON_RESIDUAL
#define POISSON_RESIDUAL

#include "domain.h"

namespace Poisson {
namespace detail {
template <class T>
inline void refresh_boundaries(Domain<T>& domain, bool fetch_neighbor) {
    // one device: Dirichlet updates are no-ops and the Neumann faces'
    // refreshes are independent, so all of them go in one launch
    bool single = true;
    for (int q = 0; q < 6; q++) single = single && !domain.face(q)->is_internal_boundary();
    if (single) {
        const mgp_array u = domain.u->desc();
        const mgp_faces F = domain.faces();
        mgpx::check(mgpx::Ops<T>::bc_neumann_faces(&u, &F, mgpx::stream()), "boundary refresh");
        return;
    }
    domain.east->update(domain, false, fetch_neighbor);
    domain.west->update(domain, false, fetch_neighbor);
    domain.north->update(domain, false, fetch_neighbor);
    domain.south->update(domain, false, fetch_neighbor);
    domain.top->update(domain, false, fetch_neighbor);
    domain.bottom->update(domain, false, fetch_neighbor);
}
}  // namespace detail

// the boundary refresh residual() performs before its kernel
template <class T>
inline void detail_refresh_for_residual(Domain<T>& domain) {
    detail::refresh_boundaries(domain, true);
}

template <class T>
void residual(Domain<T>& domain, bool fetch_neighbor) {
    detail::refresh_boundaries(domain, fetch_neighbor);
    int64_t b[6];
    domain.box(b);
    const mgp_array u = domain.u->desc(), f = domain.f->desc(), r = domain.r->desc();
    mgpx::check(mgpx::Ops<T>::residual(domain.settings.mode, &u, &f, &r, b, domain.settings.h, nullptr, mgpx::stream()),
                "residual");
}

// addition: residual plus max |r| over r's shape into *norm_dev (a device
// double the caller zeroed), in one pass
template <class T>
void residual_norm(Domain<T>& domain, double* norm_dev, bool fetch_neighbor = true) {
    detail::refresh_boundaries(domain, fetch_neighbor);
    int64_t b[6];
    domain.box(b);
    const mgp_array u = domain.u->desc(), f = domain.f->desc(), r = domain.r->desc();
    mgpx::check(mgpx::Ops<T>::residual(domain.settings.mode, &u, &f, &r, b, domain.settings.h, norm_dev, mgpx::stream()),
                "residual");
}
// addition: max |f - L u| into *norm_dev without writing r (r is written
// with the same values when the one-pass form is not available)
template <class T>
void residual_norm_only(Domain<T>& domain, double* norm_dev) {
    detail::refresh_boundaries(domain, true);
    int64_t b[6];
    domain.box(b);
    const mgp_array u = domain.u->desc(), f = domain.f->desc(), r = domain.r->desc();
    if (mgpx::Ops<T>::residual_norm(domain.settings.mode, &u, &f, &r, b, domain.settings.h, norm_dev,
                                    mgpx::stream()) == MGP_OK)
        return;
    mgpx::check(mgpx::Ops<T>::residual(domain.settings.mode, &u, &f, &r, b, domain.settings.h, norm_dev, mgpx::stream()),
                "residual");
}

// addition: the same residual on an FP64 level, written as FP32 into `out`
// (same layout) instead of r, with max |r| into *norm_dev; r is only read
// (outside the box, for the norm)
inline void residual_norm_to_f32(Domain<double>& domain, const mgp_array& out, double* norm_dev) {
    detail::refresh_boundaries(domain, true);
    int64_t b[6];
    domain.box(b);
    const mgp_array u = domain.u->desc(), f = domain.f->desc(), r = domain.r->desc();
    mgpx::check(mgp_residual_cast_f64_f32(domain.settings.mode, &u, &f, &r, &out, b, domain.settings.h, norm_dev,
                                          mgpx::stream()),
                "residual");
}
// addition: the refinement update u += (double) e and the residual of the
// updated u, written as FP32 into `out`, in one pass
// (mgp_update_residual_f64_f32); u + e goes to the second u buffer, which
// becomes u.  Returns false, with nothing done, where the domain has no second
// buffer.
inline bool update_residual_norm_to_f32(Domain<double>& domain, const mgp_array& e, const mgp_array& out,
                                        double* norm_dev) {
    if (!domain.u_alt) return false;
    int64_t b[6];
    domain.box(b);
    const mgp_array u = domain.u->desc(), f = domain.f->desc(), r = domain.r->desc();
    mgp_array un = u;
    un.data = domain.u_alt;
    const mgp_faces F = domain.faces();
    mgpx::check(mgp_update_residual_f64_f32(domain.settings.mode, &u, &un, &e, &f, &r, &out, &F, b, norm_dev,
                                            mgpx::stream()),
                "update + residual");
    domain.swap_u_alt();
    return true;
}
}  // namespace Poisson
#endif
