// Settings of one multigrid level (reference include/domainsettings.h,
// src/domainsettings.cpp:12-29): dims (n >> l) + 1, spacing h * 2^l.
#ifndef DOMAIN_SETTINGS
#define DOMAIN_SETTINGS

#include <iostream>
#include <stdexcept>
#include <string>

#include "definitions.h"
#include "settings.h"

namespace Poisson {
class DomainSettings : public Settings {
public:
    DomainSettings() : Settings() {}

    DomainSettings(const Settings& _settings, const uint_t l) : Settings(_settings) {
        if (l > 0) {
            uint_t factor = 1;
            uint_t original_dims[3] = {this->dims[0], this->dims[1], this->dims[2]};
            this->dims[0] = (this->dims[0] >> l) + 1;
            this->dims[1] = (this->dims[1] >> l) + 1;
            this->dims[2] = (this->dims[2] >> l) + 1;
            this->h *= (double_t)(factor << l);
            if ((((this->dims[0] - 1) << l) + 1 != original_dims[0]) ||
                (((this->dims[1] - 1) << l) + 1 != original_dims[1]) ||
                (((this->dims[2] - 1) << l) + 1 != original_dims[2])) {
                std::string errormsg = "Invalid grid of size: (" + std::to_string(this->dims[0]) + ",";
                errormsg += std::to_string(this->dims[1]) + "," + std::to_string(this->dims[2]) + ")";
                throw std::invalid_argument(errormsg);
            }
        }
    }

    ~DomainSettings() {}
};

inline std::ostream& operator<<(std::ostream& os, const DomainSettings& settings) {
    os << "DomainSettings:" << std::endl;
    os << "\tDomain size: (" << settings.dims[0] << "," << settings.dims[1] << "," << settings.dims[2] << ")"
       << std::endl;
    os << "\tGridSpacing: " << settings.h << std::endl;
    return os;
}
}  // namespace Poisson
#endif
