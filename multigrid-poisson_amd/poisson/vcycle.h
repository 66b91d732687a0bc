// Recursive V-cycle (reference include/vcycle.h:14-98): pre-smooth; on the
// coarsest level smoothing only; defect, restriction to the coarse right-hand
// side, zero coarse guess, coarse boundary data, recursion, prolongation and
// correction, post-smooth.
#ifndef POISSON_VCYCLE
#define POISSON_VCYCLE

#include <type_traits>

#include "gaussseidel.h"
#include "jacobi.h"
#include "residual.h"

namespace Poisson {
namespace detail {
// vcycle.h:34-76 (also fcycle.h:130-172)
template <class T>
void descend(Domain<T>*** domains, Restriction<T>& restriction, uint_t level, uint_t num_devices) {
    for (uint_t g = 0; g < num_devices; g++) {
        Domain<T>& d = *domains[g][level];
        Domain<T>& c = *domains[g][level + 1];
        d.materialize_zero();
        // consistent mode with fused kernels: residual and full weighting in
        // one pass (the fine r is not needed afterwards: the prolongation is
        // fused with the correction)
        bool done = false;
        if (d.settings.mode == CONSISTENT && d.settings.fused_sweeps && restriction.mgp_kind() == MGP_FULL_WEIGHTING) {
            detail_refresh_for_residual(d);
            int64_t b[6];
            d.box(b);
            const mgp_array u = d.u->desc(), f = d.f->desc(), r = d.r->desc(), cf = c.f->desc();
            done = mgpx::Ops<T>::residual_restrict_fw(d.settings.mode, &u, &f, &r, &cf, b, d.settings.h,
                                                      mgpx::stream()) == MGP_OK;
        }
        if (!done) {
            residual<T>(d);
            restriction.restrict_to(*d.r, *c.f);
        }
        c.u->init_zero();
        c.north->restrict_to(*d.u, *d.north, d.settings, restriction);
        c.south->restrict_to(*d.u, *d.south, d.settings, restriction);
        c.east->restrict_to(*d.u, *d.east, d.settings, restriction);
        c.west->restrict_to(*d.u, *d.west, d.settings, restriction);
        c.top->restrict_to(*d.u, *d.top, d.settings, restriction);
        c.bottom->restrict_to(*d.u, *d.bottom, d.settings, restriction);
    }
}

// addition: the coarsest level solved by CG in FP64 (mgp_cg_solve_*), in place
// on u; false where the level does not fit one workgroup (smoothing then)
template <class T>
bool coarse_cg_solve(Domain<T>& d) {
    d.materialize_zero();
    int64_t b[6];
    d.box(b);
    const mgp_array u = d.u->desc(), f = d.f->desc();
    const mgp_faces F = d.faces();
    const int rc = std::is_same<T, float>::value ? mgp_cg_solve_f32(&u, &f, &F, b, 2000, 1e-8, mgpx::stream())
                                                 : mgp_cg_solve_f64(&u, &f, &F, b, 2000, 1e-8, mgpx::stream());
    if (rc == MGP_ERR_ARG) return false;
    mgpx::check(rc, "coarsest-level CG");
    return true;
}

}  // namespace detail

// pending: u on `level` still lacks the correction P(u_{level+1}) of the
// F-cycle step before (vcycle.h:81-94, deferred by Fcycle); the first smoothing
// adds it.  The correction of this cycle's own recursion goes in with the
// post-smoothing (Relaxation::relax_prolonged: u_l += P(u_{l+1}) with r_l as the
// scratch, then the sweeps, or both in one pass)
template <class T>
void Vcycle(Domain<T>*** domains, Restriction<T>& restriction, Prolongation<T>& prolongation,
            Relaxation<T>& relaxation, T omega, uint_t level, uint_t levels, uint_t num_devices, int_t nsmooth = 10,
            bool pending = false) {
    if (level >= levels - 1) {
        Domain<T>& d = *domains[0][level];
        if (d.below) {  // coarser levels kept elsewhere: the same V-cycle through them
            relaxation.relax(domains, omega, nsmooth, level, num_devices);
            d.below->enter(d);
            d.below->vcycle(omega, nsmooth);
            d.below->leave(d);
            relaxation.relax(domains, omega, nsmooth, level, num_devices);
            return;
        }
        // coarsest level: smoothing only (vcycle.h:27-31); in CONSISTENT mode
        // optionally more passes, or a CG solve of the level (coarse_cg)
        const Settings& st = d.settings;
        if (st.mode == CONSISTENT && st.coarse_cg && detail::coarse_cg_solve(d)) return;
        const int_t n = (st.mode == CONSISTENT && st.coarse_sweeps > 0) ? st.coarse_sweeps : nsmooth;
        relaxation.relax(domains, omega, n, level, num_devices);
        return;
    }
    if (pending) relaxation.relax_prolonged(domains, prolongation, omega, nsmooth, level, num_devices);
    else relaxation.relax(domains, omega, nsmooth, level, num_devices);
    detail::descend(domains, restriction, level, num_devices);
    Vcycle<T>(domains, restriction, prolongation, relaxation, omega, level + 1, levels, num_devices, nsmooth);
    relaxation.relax_prolonged(domains, prolongation, omega, nsmooth, level, num_devices);
}
}  // namespace Poisson
#endif
