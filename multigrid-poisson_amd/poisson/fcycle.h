// F-cycle (reference include/fcycle.h:13-97): pre-smooth, descend, F on the
// next level (while two or more levels remain below), then V on the next
// level, prolong and correct, and a closing V-cycle from the finest level.
#ifndef POISSON_FCYCLE
#define POISSON_FCYCLE

#include "vcycle.h"

namespace Poisson {
// Returns true when the step's last operation, the correction u_level +=
// P(u_{level+1}), is left to the V-cycle that follows on this level (its first
// smoothing adds it, Vcycle's `pending`); level 0 runs that V-cycle itself.
template <class T>
bool Fcycle(Domain<T>*** domains, Restriction<T>& restriction, Prolongation<T>& prolongation,
            Relaxation<T>& relaxation, T omega, uint_t level, uint_t levels, uint_t num_devices, int_t nsmooth = 10) {
    // with coarser levels kept elsewhere (CoarseLevels) below the last level,
    // the recursion continues through them: the last level descends into
    // them, and every level above it has at least two levels below
    CoarseLevels<T>* below = domains[0][levels - 1]->below;
    relaxation.relax(domains, omega, nsmooth, level, num_devices);
    bool pending = false;
    if (level == levels - 1) {
        below->enter(*domains[0][level]);
        below->fcycle(omega, nsmooth);
        below->leave(*domains[0][level]);
    } else {
        detail::descend(domains, restriction, level, num_devices);
        bool next_pending = false;
        if (level < levels - 2 || below)
            next_pending =
                Fcycle(domains, restriction, prolongation, relaxation, omega, level + 1, levels, num_devices, nsmooth);
        Vcycle<T>(domains, restriction, prolongation, relaxation, omega, level + 1, levels, num_devices, nsmooth,
                  next_pending);
        pending = true;
    }
    if (level == 0) {
        Vcycle<T>(domains, restriction, prolongation, relaxation, omega, 0, levels, num_devices, nsmooth, pending);
        pending = false;
    }
    return pending;
}
}  // namespace Poisson
#endif
