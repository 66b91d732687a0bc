// Injection restriction u_c(i,j,k) = u_f(2i,2j,2k) (reference include/injection.h:29-67).
#ifndef POISSON_INJECTION
#define POISSON_INJECTION

#include "restriction.h"

namespace Poisson {
template <class T>
class Injection : public Restriction<T> {
public:
    Injection() : Restriction<T>() {}
    ~Injection() {}
    void restrict_to(DeviceArray<T>& uin, DeviceArray<T>& uout) override {
        const mgp_array a = uin.desc(), b = uout.desc();
        mgpx::check(mgpx::Ops<T>::inject(&a, &b, mgpx::stream()), "Restriction is incompatible with injection");
    }
    int mgp_kind() const override { return MGP_INJECTION; }
};
}  // namespace Poisson
#endif
