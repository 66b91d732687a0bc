// Restriction interface (reference include/restriction.h:6-25).
#ifndef POISSON_RESTRICTION
#define POISSON_RESTRICTION

#include "devicearray.h"

namespace Poisson {
template <class T>
class Restriction {
public:
    Restriction() {}
    virtual ~Restriction() {}
    virtual void restrict_to(DeviceArray<T>& uin, DeviceArray<T>& uout) = 0;
    // addition: which built-in operator this is (MGP_INJECTION,
    // MGP_FULL_WEIGHTING), or -1 for a user-defined one
    virtual int mgp_kind() const { return -1; }
};
}  // namespace Poisson
#endif
