// Boundary-condition interface (reference include/boundary.h:12-111).
#ifndef POISSON_BOUNDARY
#define POISSON_BOUNDARY

#include "definitions.h"
#include "devicearray.h"
#include "halo.h"
#include "problem_definition.h"
#include "restriction.h"
#include "settings.h"

namespace Poisson {
typedef enum { NORTH, SOUTH, EAST, WEST, TOP, BOTTOM } Location_t;

typedef enum { DIRICHLET, NEUMANN } Boundary_t;

inline bool requires_halo(Boundary_t bound, int idx, int num_devices) {
    if ((idx < num_devices - 1) || (idx > 0)) return true;
    return bound != DIRICHLET;
}

typedef struct BoundaryCondition {
    Boundary_t north;
    Boundary_t south;
    Boundary_t east;
    Boundary_t west;
    Boundary_t top;
    Boundary_t bottom;
} BoundaryCondition;

template <typename T>
class Domain;

template <class T>
class Boundary {
protected:
    Boundary<T>* neighbor = nullptr;  // internal (device-to-device) faces only
    Location_t location;
    void init_by_fun(funptr fun, Settings& settings);
    // addition: the device copy of arr is known to be all zero (set where
    // this class zeroes it, cleared where anything else may write it); a
    // coarse face whose data is zero at every descent is then filled once,
    // not at every descent
    bool zero_on_device = false;
    void zero_once() {
        if (zero_on_device && this->arr.is_on_device()) return;
        this->arr.init_zero();
        zero_on_device = this->arr.is_on_device();
    }

public:
    DeviceArray<T> arr;

    Boundary(int_t device, Location_t location, uint_t i, uint_t j, uint_t k);
    virtual ~Boundary() {}

    virtual void init(funptr ufun, funptr dudxfun, funptr dudyfun, funptr dudzfun, Settings& settings) = 0;
    virtual void write_to(DeviceArray<T>& uarr, Settings& settings) = 0;
    virtual void update(Domain<T>& domain, bool previous = false, bool fetch_neighbor = true) = 0;
    virtual void restrict_to(DeviceArray<T>& u, Boundary<T>& boundary, Settings& settings,
                             Restriction<T>& restriction) = 0;
    virtual bool is_internal_boundary() = 0;
    virtual void init_zero() = 0;
    virtual void to_device() = 0;
    virtual void link(Boundary<T>* _boundary) = 0;
    virtual bool is_non_eliminated() = 0;

    Location_t get_location() const { return location; }  // addition
    // addition: a coarse level's face data as every descent's restrict_to
    // leaves it where that is always zero (Dirichlet faces; Neumann faces in
    // consistent mode), set before the first descent so that recorded cycles
    // need no fill at all
    void prepare_coarse(const Settings& s) {
        if (is_internal_boundary()) return;
        if (is_non_eliminated() || s.mode == CONSISTENT) zero_once();
    }
};

template <class T>
Boundary<T>::Boundary(int_t device, Location_t location, uint_t i, uint_t j, uint_t k) : arr(device, i, j, k) {
    this->location = location;
}

// Face values from an analytic function at x0 + (i + off) h etc., with the
// face's offset along its normal (boundary.h:82-111)
template <class T>
void Boundary<T>::init_by_fun(funptr fun, Settings& settings) {
    const double_t x0 = settings.origin[0];
    const double_t y0 = settings.origin[1];
    const double_t z0 = settings.origin[2];
    const double_t h = settings.h;
    double_t offx = 0, offy = 0, offz = 0;
    switch (this->location) {
    case TOP: offz = settings.dims[2] - 1; break;
    case NORTH: offy = settings.dims[1] - 1; break;
    case EAST: offx = settings.dims[0] - 1; break;
    default: break;
    }
    for (int_t i = 0; i < this->arr.shape[0]; i++)
        for (int_t j = 0; j < this->arr.shape[1]; j++)
            for (int_t k = 0; k < this->arr.shape[2]; k++)
                this->arr.at[this->arr.idx(i, j, k)] =
                    fun(x0 + (((double_t)i) + offx) * h, y0 + (((double_t)j) + offy) * h, z0 + (((double_t)k) + offz) * h);
}
}  // namespace Poisson
#endif
