// The level hierarchy's settings (reference include/grid.h, src/grid.cpp:6-13).
#ifndef POISSON_GRID
#define POISSON_GRID

#include <iostream>

#include "domainsettings.h"
#include "settings.h"

namespace Poisson {
class Grid {
public:
    const uint_t levels;
    DomainSettings* domainsettings;

    Grid(const Settings& settings, const uint_t _levels, bool print = true) : levels(_levels) {
        if (_levels < 1) throw std::invalid_argument("levels must be at least 1");
        this->domainsettings = new DomainSettings[_levels];
        try {
            for (uint_t l = 0; l < _levels; l++) {
                this->domainsettings[l] = DomainSettings(settings, l);
                if (print) std::cout << this->domainsettings[l];
            }
        } catch (...) {
            delete[] this->domainsettings;
            throw;
        }
    }

    Grid(const Grid&) = delete;
    Grid& operator=(const Grid&) = delete;

    ~Grid() { delete[] this->domainsettings; }
};
}  // namespace Poisson
#endif
