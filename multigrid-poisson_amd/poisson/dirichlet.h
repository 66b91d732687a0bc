// Dirichlet faces (reference include/dirichlet.h:7-134): values written onto
// the boundary nodes once; boundary nodes are outside the relaxation box;
// homogeneous on coarse levels.
#ifndef DIRICHLET_BOUNDARY
#define DIRICHLET_BOUNDARY

#include "boundary.h"

namespace Poisson {
template <class T>
class Dirichlet : public Boundary<T> {
public:
    Dirichlet(int_t device, Location_t location, uint_t i, uint_t j, uint_t k) : Boundary<T>(device, location, i, j, k) {}
    virtual ~Dirichlet() {}

    void init(funptr ufun, funptr dudxfun, funptr dudyfun, funptr dudzfun, Settings& settings) override {
        (void)dudxfun, (void)dudyfun, (void)dudzfun;
        this->init_by_fun(ufun, settings);
    }
    void write_to(DeviceArray<T>& uarr, Settings& settings) override {
        (void)settings;
        const mgp_array u = uarr.desc();
        mgpx::check(mgpx::Ops<T>::bc_dirichlet(&u, (int)this->location, this->arr.devptr, mgpx::stream()),
                    "Dirichlet::write_to");
    }
    // the values are written once at to_device (dirichlet.h:95-99)
    void update(Domain<T>& domain, bool previous, bool fetch_neighbor) override {
        (void)domain, (void)previous, (void)fetch_neighbor;
    }
    // the coarse-grid error vanishes on a Dirichlet face (dirichlet.h:101-107)
    void restrict_to(DeviceArray<T>& u, Boundary<T>& boundary, Settings& settings,
                     Restriction<T>& restriction) override {
        (void)u, (void)boundary, (void)settings, (void)restriction;
        this->zero_once();
    }
    bool is_internal_boundary() override { return false; }
    void init_zero() override {
        this->arr.init_zero();
        this->zero_on_device = this->arr.is_on_device();
    }
    void to_device() override {
        this->arr.to_device();
        this->zero_on_device = false;
    }
    void link(Boundary<T>* _boundary) override { (void)_boundary; }
    bool is_non_eliminated() override { return true; }
};
}  // namespace Poisson
#endif
