// Smoother driver (reference include/relaxation.h:7-94).
#ifndef POISSON_RELAXATION
#define POISSON_RELAXATION

#include "domain.h"

namespace Poisson {
template <class T>
class Relaxation {
public:
    virtual ~Relaxation() {}
    // smoothing_multiplier()*nsmooth passes on `level`, each preceded by the
    // boundary refresh (relaxation.h:19-94)
    void relax(Domain<T>*** domain, T omega, int_t nsmooth, int_t level, int_t num_devices);
    virtual T default_omega() = 0;
    virtual bool requires_duplicate_solution() = 0;
    virtual int smoothing_multiplier() = 0;
    // multi-device send buffers (relaxation.h:96-132): no internal faces on one GPU
    void fill_send_buffer(Domain<T>& domain, const T omega, Location_t loc) { (void)domain, (void)omega, (void)loc; }
    virtual void relaxation_kernel(Domain<T>& domain, T omega, const int_t xmin, const int_t xmax, const int_t ymin,
                                   const int_t ymax, const int_t zmin, const int_t zmax) = 0;

    // ---- additions: per-level pass loop (overridable), and host-side state a
    // replayed graph must restore
    virtual void relax_level(Domain<T>& d, T omega, int_t nsmooth);
    // the cycle's correction u += P(u_{level+1}) (include/vcycle.h:81-94) and
    // then relax(): a smoother may add the correction inside its first pass
    void relax_prolonged(Domain<T>*** domains, Prolongation<T>& prolongation, T omega, int_t nsmooth, int_t level,
                         int_t num_devices);
    // one level: prolong_add into d.u (d.r as the scratch), then relax_level
    virtual void relax_level_prolonged(Domain<T>& d, Domain<T>& coarse, Prolongation<T>& prolongation, T omega,
                                       int_t nsmooth) {
        d.materialize_zero();
        prolongation.prolong_add(*coarse.u, *d.u, *d.r);
        relax_level(d, omega, nsmooth);
    }
    virtual int mgp_state() const { return 0; }
    virtual void mgp_set_state(int) {}
};

template <class T>
void Relaxation<T>::relax(Domain<T>*** domains, T omega, int_t nsmooth, int_t level, int_t num_devices) {
    for (int_t gpuid = 0; gpuid < num_devices; gpuid++) this->relax_level(*domains[gpuid][level], omega, nsmooth);
}

template <class T>
void Relaxation<T>::relax_prolonged(Domain<T>*** domains, Prolongation<T>& prolongation, T omega, int_t nsmooth,
                                    int_t level, int_t num_devices) {
    if (num_devices == 1) {
        this->relax_level_prolonged(*domains[0][level], *domains[0][level + 1], prolongation, omega, nsmooth);
        return;
    }
    for (int_t g = 0; g < num_devices; g++) {
        domains[g][level]->materialize_zero();
        prolongation.prolong_add(*domains[g][level + 1]->u, *domains[g][level]->u, *domains[g][level]->r);
    }
    relax(domains, omega, nsmooth, level, num_devices);
}

template <class T>
void Relaxation<T>::relax_level(Domain<T>& domain, T omega, int_t nsmooth) {
    domain.materialize_zero();
    const bool dup = requires_duplicate_solution();
    for (int_t s = 0; s < smoothing_multiplier() * nsmooth; s++) {
        if (dup) domain.swap_u();
        domain.east->update(domain, dup);
        domain.west->update(domain, dup);
        domain.north->update(domain, dup);
        domain.south->update(domain, dup);
        domain.top->update(domain, dup);
        domain.bottom->update(domain, dup);
        int64_t b[6];
        domain.box(b);
        relaxation_kernel(domain, omega, b[0], b[1], b[2], b[3], b[4], b[5]);
    }
}
}  // namespace Poisson
#endif
