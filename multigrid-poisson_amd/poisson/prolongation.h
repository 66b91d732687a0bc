// Prolongation interface (reference include/prolongation.h:6-25).
#ifndef POISSON_PROLONGATION
#define POISSON_PROLONGATION

#include "devicearray.h"

namespace Poisson {
template <class T>
class Prolongation {
public:
    Prolongation() {}
    virtual ~Prolongation() {}
    virtual void prolong(DeviceArray<T>& uin, DeviceArray<T>& uout) = 0;
    // addition: u += P(coarse).  The cycles call this where the reference
    // prolongs into the level's r and then adds r to u (vcycle.h:81-94).
    // The default does exactly that; TrilinearInterpolation fuses both into one
    // pass.  Either way r's points outside the relaxation box end up zero,
    // which is what the reference's prolongation leaves there (coarse boundary
    // values are zero), so every later reader sees the same data.
    virtual void prolong_add(DeviceArray<T>& coarse, DeviceArray<T>& u, DeviceArray<T>& scratch) {
        this->prolong(coarse, scratch);
        u.add(scratch);
    }
    // addition: true for the trilinear prolongation, which smoothers may fuse
    // into their first pass (mgp_rbgs_sweep2_prolong_*)
    virtual bool mgp_trilinear() const { return false; }
};
}  // namespace Poisson
#endif
