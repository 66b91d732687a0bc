// Trilinear prolongation (reference include/trilinearinterpolation.h:29-134):
// one kernel whose values are bit-identical to the reference's four passes.
#ifndef POISSON_TRILINEAR_INTERPOLATION
#define POISSON_TRILINEAR_INTERPOLATION

#include "prolongation.h"

namespace Poisson {
template <class T>
class TrilinearInterpolation : public Prolongation<T> {
public:
    TrilinearInterpolation() : Prolongation<T>() {}
    ~TrilinearInterpolation() {}
    void prolong(DeviceArray<T>& uin, DeviceArray<T>& uout) override {
        const mgp_array a = uin.desc(), b = uout.desc();
        mgpx::check(mgpx::Ops<T>::prolong(&a, &b, mgpx::stream()),
                    "Prolongation is incompatible with trilinear interpolation");
    }
    bool mgp_trilinear() const override { return true; }
    void prolong_add(DeviceArray<T>& coarse, DeviceArray<T>& u, DeviceArray<T>& scratch) override {
        (void)scratch;
        const mgp_array a = coarse.desc(), b = u.desc();
        mgpx::check(mgpx::Ops<T>::prolong_add(&a, &b, mgpx::stream()),
                    "Prolongation is incompatible with trilinear interpolation");
    }
};
}  // namespace Poisson
#endif
