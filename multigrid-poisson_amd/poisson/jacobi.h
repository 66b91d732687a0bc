// Weighted Jacobi, default omega 6/7 (reference include/jacobi.h:16-82).
#ifndef POISSON_JACOBI
#define POISSON_JACOBI

#include "relaxation.h"

namespace Poisson {
template <class T>
class Jacobi : public Relaxation<T> {
public:
    T default_omega() override { return 6.0 / 7.0; }
    bool requires_duplicate_solution() override { return true; }
    int smoothing_multiplier() override { return 1; }
    // u = (1-w) uprev + w/6 (sum_nb uprev - h^2 f) over the box
    void relaxation_kernel(Domain<T>& domain, T omega, const int_t xmin, const int_t xmax, const int_t ymin,
                           const int_t ymax, const int_t zmin, const int_t zmax) override {
        const int64_t b[6] = {xmin, xmax, ymin, ymax, zmin, zmax};
        const mgp_array u = domain.u->desc(), v = domain.uprev->desc(), f = domain.f->desc();
        mgpx::check(mgpx::Ops<T>::jacobi(domain.settings.mode, &u, &v, &f, b, (double)omega, domain.settings.h,
                                         mgpx::stream()),
                    "Jacobi::relaxation_kernel");
    }
};
}  // namespace Poisson
#endif
