// Settings of one multigrid level (the reference's DomainSettings:
// include/domainsettings.h, src/domainsettings.cpp:12-29): the coarsening rule
// dims (n >> l) + 1, spacing h 2^l, and the error text are the reference's
// interface; the code is this library's.
#ifndef DOMAIN_SETTINGS
#define DOMAIN_SETTINGS

#include <cmath>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "definitions.h"
#include "settings.h"

namespace Poisson {
class DomainSettings : public Settings {
public:
    DomainSettings() : Settings() {}

    // level l of a hierarchy over `fine`: every axis keeps (n >> l) + 1 points
    // and the spacing grows to h 2^l (exact: a power-of-two scale); the fine
    // axis must hold 2^l m + 1 points so that each coarse point sits on a
    // fine one, else std::invalid_argument names the coarse size
    DomainSettings(const Settings& fine, const uint_t l) : Settings(fine) {
        if (l == 0) return;
        const uint_t step = (uint_t)1 << l;
        bool nested = true;
        for (int axis = 0; axis < 3; axis++) {
            nested = nested && fine.dims[axis] >= 1 && (fine.dims[axis] - 1) % step == 0;
            dims[axis] = (fine.dims[axis] >> l) + 1;
        }
        h = std::ldexp(fine.h, (int)l);
        if (!nested) {
            std::ostringstream msg;
            msg << "Invalid grid of size: (" << dims[0] << "," << dims[1] << "," << dims[2] << ")";
            throw std::invalid_argument(msg.str());
        }
    }

    ~DomainSettings() {}
};

// the level listing a solver prints at construction (the reference's text)
inline std::ostream& operator<<(std::ostream& os, const DomainSettings& level) {
    const uint_t* n = level.dims;
    return os << "DomainSettings:\n\tDomain size: (" << n[0] << "," << n[1] << "," << n[2] << ")\n"
              << "\tGridSpacing: " << level.h << std::endl;
}
}  // namespace Poisson
#endif
