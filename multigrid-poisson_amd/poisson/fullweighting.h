// 27-point full-weighting restriction with clamped edges
// (reference include/fullweighting.h:37-122).
#ifndef POISSON_FULL_WEIGHT
#define POISSON_FULL_WEIGHT

#include "restriction.h"

namespace Poisson {
template <class T>
class FullWeighting : public Restriction<T> {
public:
    FullWeighting() : Restriction<T>() {}
    ~FullWeighting() {}
    void restrict_to(DeviceArray<T>& uin, DeviceArray<T>& uout) override {
        const mgp_array a = uin.desc(), b = uout.desc();
        mgpx::check(mgpx::Ops<T>::fw(&a, &b, mgpx::stream()), "Restriction is incompatible with full weighting");
    }
    int mgp_kind() const override { return MGP_FULL_WEIGHTING; }
};
}  // namespace Poisson
#endif
