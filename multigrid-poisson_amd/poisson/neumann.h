// Neumann faces (reference include/neumann.h:7-280): ghost = sign*2h*g +
// u(mirror), refreshed before every sweep and residual; boundary nodes are
// unknowns.  Coarse-level data: the reference's one-sided-difference defect
// (SolverMode REFERENCE) or zero, the homogeneous correction (CONSISTENT).
#ifndef NEUMANN_BOUNDARY
#define NEUMANN_BOUNDARY

#include "boundary.h"

namespace Poisson {
template <class T>
class Neumann : public Boundary<T> {
public:
    Neumann(int_t device, Location_t location, uint_t i, uint_t j, uint_t k) : Boundary<T>(device, location, i, j, k) {}
    virtual ~Neumann() {}

    void init(funptr ufun, funptr dudxfun, funptr dudyfun, funptr dudzfun, Settings& settings) override {
        (void)ufun;
        switch (this->location) {
        case EAST:
        case WEST: this->init_by_fun(dudxfun, settings); break;
        case NORTH:
        case SOUTH: this->init_by_fun(dudyfun, settings); break;
        default: this->init_by_fun(dudzfun, settings); break;
        }
    }
    void write_to(DeviceArray<T>& uarr, Settings& settings) override {
        const mgp_array u = uarr.desc();
        mgpx::check(mgpx::Ops<T>::bc_neumann(&u, (int)this->location, this->arr.devptr, settings.h, mgpx::stream()),
                    "Neumann::write_to");
    }
    void update(Domain<T>& domain, bool previous, bool fetch_neighbor) override;
    void restrict_to(DeviceArray<T>& u, Boundary<T>& boundary, Settings& settings,
                     Restriction<T>& restriction) override {
        if (settings.mode == CONSISTENT) {
            this->zero_once();  // homogeneous coarse data
            return;
        }
        this->zero_on_device = false;
        const int kind = restriction.mgp_kind();
        if (kind < 0)
            throw std::invalid_argument("Neumann::restrict_to: reference-mode coarse Neumann data needs a built-in restriction");
        const mgp_array uf = u.desc();
        mgpx::check(mgpx::Ops<T>::bc_restrict_neumann(kind, &uf, (int)this->location, boundary.arr.devptr,
                                                      this->arr.devptr, settings.h, mgpx::stream()),
                    "Neumann::restrict_to");
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
    bool is_non_eliminated() override { return false; }
};
}  // namespace Poisson
#endif
