// Red-black Gauss-Seidel SOR, default omega 8/9, one colour per
// relaxation_kernel call starting with red (reference include/gaussseidel.h:14-99).
//
// Where the colouring is a checkerboard (always in CONSISTENT mode; in
// REFERENCE mode when both outer strides are odd) and Settings::fused_sweeps is
// on, relax_level runs each red+black pair as one out-of-place pass
// (mgp_rbgs_sweep) that forms Neumann ghosts on the fly; the values are
// identical to the two half-sweeps with boundary refreshes in between; pairs
// of such sweeps run as one pass (mgp_rbgs_sweep2) unless the "sweep.pair"
// tuning key is 0 or the level has fewer points than "sweep.pair_min_points".
#ifndef POISSON_GAUSS_SEIDEL
#define POISSON_GAUSS_SEIDEL

#include <type_traits>

#include "relaxation.h"

namespace Poisson {
template <class T>
class GaussSeidel : public Relaxation<T> {
private:
    bool alternate = true;

public:
    T default_omega() override { return 8.0 / 9.0; }
    int smoothing_multiplier() override { return 2; }
    bool requires_duplicate_solution() override { return false; }

    void relaxation_kernel(Domain<T>& domain, T omega, const int_t xmin, const int_t xmax, const int_t ymin,
                           const int_t ymax, const int_t zmin, const int_t zmax) override {
        const int64_t b[6] = {xmin, xmax, ymin, ymax, zmin, zmax};
        const mgp_array u = domain.u->desc(), f = domain.f->desc();
        mgpx::check(mgpx::Ops<T>::rbgs_half(domain.settings.mode, &u, &f, b, (double)omega, domain.settings.h,
                                            alternate ? 0 : 1, mgpx::stream()),
                    "GaussSeidel::relaxation_kernel");
        alternate = !alternate;
    }

    static bool checkerboard(const Domain<T>& d) {
        return d.settings.mode == CONSISTENT || ((d.u->stride[0] & 1) && (d.u->stride[1] & 1));
    }

    void relax_level(Domain<T>& d, T omega, int_t nsmooth) override {
        if (d.u_zero && from_zero(d, omega, nsmooth)) return;
        d.materialize_zero();
        if (!d.u_alt || !d.settings.fused_sweeps || !checkerboard(d) || !alternate || nsmooth < 1) {
            Relaxation<T>::relax_level(d, omega, nsmooth);
            return;
        }
        int64_t b[6];
        d.box(b);
        const mgp_faces F = d.faces();
        const mgp_array f = d.f->desc();
        if (d.settings.fused_sweeps) {
            // a level that fits one workgroup's LDS: every pass in one launch
            // (the result lands where the last of nsmooth ping-pong passes would)
            const mgp_array ui = d.u->desc();
            mgp_array uo = ui;
            if (nsmooth & 1) uo.data = d.u_alt;
            // FP32 coarsest level of a refinement solve: the passes in FP64
            // arithmetic where the level fits the LDS in doubles
            const bool dp = std::is_same<T, float>::value && d.settings.coarse_fp64 && d.coarsest && !d.below;
            if (dp && mgp_rbgs_sweeps_f32_f64(d.settings.mode, &ui, &uo, &f, &F, b, (double)omega, nsmooth,
                                              mgpx::stream()) == MGP_OK) {
                if (nsmooth & 1) d.swap_u_alt();
                return;
            }
            if (mgpx::Ops<T>::rbgs_sweeps(d.settings.mode, &ui, &uo, &f, &F, b, (double)omega, nsmooth,
                                          mgpx::stream()) == MGP_OK) {
                if (nsmooth & 1) d.swap_u_alt();
                return;
            }
        }
        int_t s = 0;
        int64_t pair = 1, pair_min = 0;
        if (mgp_tuning_get("sweep.pair", &pair) != MGP_OK) pair = 1;
        // pairs pay on large levels only (two one-pass sweeps are faster below)
        if (mgp_tuning_get("sweep.pair_min_points", &pair_min) == MGP_OK &&
            d.u->shape[0] * d.u->shape[1] * d.u->shape[2] < pair_min)
            pair = 0;
        // pairs of sweeps in one pass each (MGP_ERR_ARG: a shape the two-sweep
        // pass does not take; single sweeps then)
        for (; pair && s + 2 <= nsmooth; s += 2) {
            mgp_array ui = d.u->desc();
            mgp_array uo = ui;
            uo.data = d.u_alt;
            const int rc = mgpx::Ops<T>::rbgs_sweep2(d.settings.mode, &ui, &uo, &f, &F, b, (double)omega,
                                                     mgpx::stream());
            if (rc == MGP_ERR_ARG) break;
            mgpx::check(rc, "GaussSeidel sweep pair");
            d.swap_u_alt();
        }
        for (; s < nsmooth; s++) {
            mgp_array ui = d.u->desc();
            mgp_array uo = ui;
            uo.data = d.u_alt;
            mgpx::check(mgpx::Ops<T>::rbgs_sweep(d.settings.mode, &ui, &uo, &f, &F, b, (double)omega, mgpx::stream()),
                        "GaussSeidel sweep");
            d.swap_u_alt();
        }
    }

    // u zero in value, its buffer unwritten (Domain::u_zero): where the first
    // pair of sweeps runs through the two-sweep pass, that pass starts from
    // zeros without reading u (mgp_rbgs_sweep2_from_zero_*); false where it does
    // not apply (the caller then writes the zeros and smooths as usual)
    bool from_zero(Domain<T>& d, T omega, int_t nsmooth) {
        int64_t pair = 1, pair_min = 0;
        if (mgp_tuning_get("sweep.pair", &pair) != MGP_OK) pair = 1;
        if (mgp_tuning_get("sweep.pair_min_points", &pair_min) != MGP_OK) pair_min = 0;
        if (!(pair && d.u_alt && d.settings.fused_sweeps && d.settings.mode == CONSISTENT && alternate &&
              nsmooth >= 2 && d.u->shape[0] * d.u->shape[1] * d.u->shape[2] >= pair_min))
            return false;
        int64_t b[6];
        d.box(b);
        const mgp_faces F = d.faces();
        const mgp_array f = d.f->desc();
        mgp_array uo = d.u->desc();
        uo.data = d.u_alt;
        const int rc = mgpx::Ops<T>::rbgs_sweep2_from_zero(d.settings.mode, &uo, &f, &F, b, (double)omega,
                                                           mgpx::stream());
        if (rc == MGP_ERR_ARG) return false;
        mgpx::check(rc, "GaussSeidel sweep pair from zero");
        d.swap_u_alt();
        d.u_zero = false;
        if (nsmooth > 2) relax_level(d, omega, nsmooth - 2);
        return true;
    }

    // the correction and the first pair of sweeps in one pass
    // (mgp_rbgs_sweep2_prolong_*) where relax_level would run that pair through
    // the two-sweep pass: consistent mode, trilinear prolongation, a level of at
    // least "sweep.pair_min_points" points, "sweep2.prolong" 2 (FP64 and FP32
    // levels, default) or 1 (FP64 levels only).  1025 x 1025 x 513: FP64 5.18 ms
    // against 4.54 + 2.05 ms for the separate passes, FP32 2.43 against
    // 2.30 + 1.01 ms
    void relax_level_prolonged(Domain<T>& d, Domain<T>& coarse, Prolongation<T>& prolongation, T omega,
                               int_t nsmooth) override {
        d.materialize_zero();
        int64_t pair = 1, pair_min = 0, fuse = 2;
        if (mgp_tuning_get("sweep.pair", &pair) != MGP_OK) pair = 1;
        if (mgp_tuning_get("sweep.pair_min_points", &pair_min) != MGP_OK) pair_min = 0;
        if (mgp_tuning_get("sweep2.prolong", &fuse) != MGP_OK) fuse = 2;
        if (!std::is_same<T, double>::value && fuse < 2) fuse = 0;
        const bool applies = fuse && pair && d.u_alt && d.settings.fused_sweeps && d.settings.mode == CONSISTENT &&
                             alternate && nsmooth >= 2 && prolongation.mgp_trilinear() &&
                             d.u->shape[0] * d.u->shape[1] * d.u->shape[2] >= pair_min;
        if (applies) {
            int64_t b[6];
            d.box(b);
            const mgp_faces F = d.faces();
            const mgp_array f = d.f->desc(), ui = d.u->desc(), c = coarse.u->desc();
            mgp_array uo = ui;
            uo.data = d.u_alt;
            const int rc = mgpx::Ops<T>::rbgs_sweep2_prolong(d.settings.mode, &ui, &uo, &f, &F, b, (double)omega, &c,
                                                             mgpx::stream());
            if (rc == MGP_OK) {
                d.swap_u_alt();
                if (nsmooth > 2) relax_level(d, omega, nsmooth - 2);
                return;
            }
            if (rc != MGP_ERR_ARG) mgpx::check(rc, "GaussSeidel sweep pair after the prolongation");
        }
        Relaxation<T>::relax_level_prolonged(d, coarse, prolongation, omega, nsmooth);
    }

    int mgp_state() const override { return alternate ? 1 : 0; }
    void mgp_set_state(int s) override { alternate = s != 0; }
};
}  // namespace Poisson
#endif
