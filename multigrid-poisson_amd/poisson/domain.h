// One multigrid level on one GPU (reference include/domain.h:61-344): the
// arrays u, uprev (Jacobi only; Gauss-Seidel aliases it to u), f and r with
// halos on the faces that need ghosts, and the six boundary objects.
#ifndef TOP3D_DOMAIN
#define TOP3D_DOMAIN

#include <stdexcept>
#include <string>
#include <utility>

#include "boundary.h"
#include "definitions.h"
#include "dirichlet.h"
#include "halo.h"
#include "neumann.h"
#include "problem_definition.h"
#include "prolongation.h"
#include "restriction.h"
#include "settings.h"

namespace Poisson {
template <typename T>
class Relaxation;
template <typename T>
class Jacobi;
template <typename T>
class GaussSeidel;
template <typename T>
void residual(Domain<T>& domain, bool fetch_neighbor = true);

// addition: the levels below a domain when another object keeps them (the
// FP16 coarse levels of an FP32 cycle, fp16levels.h).  enter() forms the
// domain's residual and restricts it into the first of them with a zero
// guess; vcycle() / fcycle() run what the cycle templates run on those
// levels (fcycle: the F-cycle's recursion below the domain, F where two or
// more levels remain, then V); leave() adds the prolonged correction to the
// domain's u.
template <class T>
class CoarseLevels {
public:
    virtual ~CoarseLevels() {}
    virtual void enter(Domain<T>& d) = 0;
    virtual void vcycle(T omega, int_t nsmooth) = 0;
    virtual void fcycle(T omega, int_t nsmooth) = 0;
    virtual void leave(Domain<T>& d) = 0;
};

template <class T>
class Domain {
private:
    Halo halo;
    friend class Relaxation<T>;
    friend class Jacobi<T>;
    friend class GaussSeidel<T>;
    friend void residual<T>(Domain<T>& domain, bool fetch_neighbor);
    void init_f(funptr ffun);
    bool is_initialized;
    const bool requires_duplicate;

protected:
    void write_bc_to(DeviceArray<T>& uarr);

public:
    Settings settings;
    DeviceArray<T>* u;
    DeviceArray<T>* uprev;
    DeviceArray<T>* f;
    DeviceArray<T>* r;
    Boundary<T>* north;
    Boundary<T>* south;
    Boundary<T>* east;
    Boundary<T>* west;
    Boundary<T>* top;
    Boundary<T>* bottom;

    // addition: second device buffer of u for out-of-place red-black sweeps
    // (Gauss-Seidel with fused sweeps only; nullptr otherwise)
    T* u_alt = nullptr;
    // addition: coarser levels kept elsewhere (the last level of a hierarchy
    // only; nullptr otherwise)
    CoarseLevels<T>* below = nullptr;
    // addition: the last level of its solver's hierarchy
    bool coarsest = false;
    // addition: u is zero in value though its buffer was not written (the
    // correction of an iterative-refinement step before its cycle); the first
    // smoothing takes it as zero without reading it (mgp_rbgs_sweep2_from_zero_*),
    // and any other use of u first writes the zeros (materialize_zero)
    bool u_zero = false;
    void materialize_zero() {
        if (!u_zero) return;
        u->init_zero();
        u_zero = false;
    }

    Domain() : halo(), is_initialized(false), requires_duplicate(true), settings() {}
    Domain(Settings& _settings, BoundaryCondition& BC, bool requires_duplicate, int_t device[3], int_t device_shape[3]);
    Domain(const Domain&) = delete;
    Domain& operator=(const Domain&) = delete;
    ~Domain();

    void init(funptr ufun, funptr ffun, funptr dudxfun, funptr dudyfun, funptr dudzfun);
    void init_zero();
    void to_device();
    // the device state to_device() gives right after init_zero(), made on the
    // device (no host transfer; the host mirrors stay untouched)
    void to_device_zero();
    void to_host();
    void save(std::string ufilename = "u.vtk");
    void save_all(std::string ufilename = "u.vtk", std::string ffilename = "f.vtk", std::string rfilename = "u.vtk");
    void save_halo();
    void save_restrict_prolong(Restriction<T>& restriction, Prolongation<T>& prolongation);
    void swap_u();

    // ---- additions
    Boundary<T>* face(int loc) const {
        Boundary<T>* b[6] = {north, south, east, west, top, bottom};
        return b[loc];
    }
    // relaxation/residual box {xmin,xmax,ymin,ymax,zmin,zmax} (relaxation.h:78-83)
    void box(int64_t b[6]) const {
        b[0] = west->is_non_eliminated();
        b[1] = u->shape[0] - east->is_non_eliminated();
        b[2] = south->is_non_eliminated();
        b[3] = u->shape[1] - north->is_non_eliminated();
        b[4] = bottom->is_non_eliminated();
        b[5] = u->shape[2] - top->is_non_eliminated();
    }
    mgp_faces faces() const {
        mgp_faces F;
        for (int q = 0; q < 6; q++) {
            Boundary<T>* bq = face(q);
            F.type[q] = bq->is_non_eliminated() ? MGP_DIRICHLET : MGP_NEUMANN;
            F.g[q] = F.type[q] == MGP_NEUMANN ? bq->arr.devptr : nullptr;
        }
        F.h = settings.h;
        return F;
    }
    void swap_u_alt() { std::swap(u->devptr, u_alt); }
    // a coarse level's face data before the first descent (Boundary::prepare_coarse)
    void prepare_coarse_faces() {
        for (int q = 0; q < 6; q++) face(q)->prepare_coarse(settings);
    }
    // u (and uprev) = 0 with the boundary data written again, on the device;
    // the state to_device() leaves for a fresh solve with unchanged f and faces
    void reset_solution() {
        u_zero = false;
        u->init_zero();
        write_bc_to(*u);
        if (requires_duplicate) {
            uprev->init_zero();
            write_bc_to(*uprev);
        }
        if (u_alt)
            mgpx::check(mgp_memcpy_d2d(u_alt, u->devptr, (size_t)u->size * sizeof(T), mgpx::stream()),
                        "Domain::reset_solution");
    }
};

template <class T>
Domain<T>::Domain(Settings& _settings, BoundaryCondition& BC, bool duplicate, int_t dev[3], int_t dev_shape[3])
    : halo(BC.east != DIRICHLET || dev[0] < dev_shape[0] - 1, BC.west != DIRICHLET || dev[0] > 0,
           BC.north != DIRICHLET || dev[1] < dev_shape[1] - 1, BC.south != DIRICHLET || dev[1] > 0,
           BC.top != DIRICHLET || dev[2] < dev_shape[2] - 1, BC.bottom != DIRICHLET || dev[2] > 0),
      is_initialized(true),
      requires_duplicate(duplicate),
      settings(_settings) {
    if (dev_shape[0] * dev_shape[1] * dev_shape[2] != 1)
        throw std::invalid_argument("Domain: one device only (multi-device decomposition is not built on one GPU)");
    this->u = new DeviceArray<T>(settings, halo);
    this->uprev = requires_duplicate ? new DeviceArray<T>(settings, halo) : this->u;
    this->f = new DeviceArray<T>(settings, halo);
    this->r = new DeviceArray<T>(settings, halo);
    const int_t dv = settings.dev;
    const uint_t nx = settings.dims[0], ny = settings.dims[1], nz = settings.dims[2];
    // the requested type per face; the reference always builds north, south
    // and bottom as Dirichlet (domain.h:139-145, 191-192) while their halos
    // still follow the request
    const bool ref = settings.mode == REFERENCE;
    auto make = [&](Boundary_t t, Location_t loc, uint_t i, uint_t j, uint_t k) -> Boundary<T>* {
        if (t == DIRICHLET) return new Dirichlet<T>(dv, loc, i, j, k);
        return new Neumann<T>(dv, loc, i, j, k);
    };
    this->top = make(BC.top, TOP, nx, ny, 1);
    this->bottom = make(ref ? DIRICHLET : BC.bottom, BOTTOM, nx, ny, 1);
    this->east = make(BC.east, EAST, 1, ny, nz);
    this->west = make(BC.west, WEST, 1, ny, nz);
    this->north = make(ref ? DIRICHLET : BC.north, NORTH, nx, 1, nz);
    this->south = make(ref ? DIRICHLET : BC.south, SOUTH, nx, 1, nz);
    if (!requires_duplicate && settings.fused_sweeps) {
        void* p = nullptr;
        mgpx::check(mgpx::dev_alloc(&p, (size_t)u->size * sizeof(T)), "Domain: u_alt");
        u_alt = static_cast<T*>(p);
    }
}

template <class T>
Domain<T>::~Domain() {
    if (is_initialized) {
        delete this->u;
        if (requires_duplicate) delete this->uprev;
        delete this->f;
        delete this->r;
        delete this->north;
        delete this->south;
        delete this->east;
        delete this->west;
        delete this->top;
        delete this->bottom;
        mgpx::dev_free(u_alt);
    }
}

// domain.h:214-227
template <class T>
void Domain<T>::init(funptr ufun, funptr ffun, funptr dudxfun, funptr dudyfun, funptr dudzfun) {
    this->north->init(ufun, dudxfun, dudyfun, dudzfun, this->settings);
    this->south->init(ufun, dudxfun, dudyfun, dudzfun, this->settings);
    this->east->init(ufun, dudxfun, dudyfun, dudzfun, this->settings);
    this->west->init(ufun, dudxfun, dudyfun, dudzfun, this->settings);
    this->top->init(ufun, dudxfun, dudyfun, dudzfun, this->settings);
    this->bottom->init(ufun, dudxfun, dudyfun, dudzfun, this->settings);
    this->init_f(ffun);
    if (requires_duplicate) this->uprev->init_zero();
    this->u->init_zero();
    this->r->init_zero();
}

template <class T>
void Domain<T>::init_zero() {
    this->north->init_zero();
    this->south->init_zero();
    this->east->init_zero();
    this->west->init_zero();
    this->top->init_zero();
    this->bottom->init_zero();
    this->f->init_zero();
    this->u->init_zero();
    if (requires_duplicate) this->uprev->init_zero();
    this->r->init_zero();
}

// domain.h:246-263: copy everything, then write the boundary data into u
// (and uprev); the out-of-place sweep buffer starts as a copy of u
template <class T>
void Domain<T>::to_device() {
    u_zero = false;
    this->u->to_device();
    if (requires_duplicate) this->uprev->to_device();
    this->f->to_device();
    this->r->to_device();
    this->north->to_device();
    this->south->to_device();
    this->east->to_device();
    this->west->to_device();
    this->top->to_device();
    this->bottom->to_device();
    this->write_bc_to(*(this->u));
    if (requires_duplicate) this->write_bc_to(*(this->uprev));
    if (u_alt)
        mgpx::check(mgp_memcpy_d2d(u_alt, u->devptr, (size_t)u->size * sizeof(T), mgpx::stream()), "Domain::to_device");
    mgpx::check(mgp_stream_sync(mgpx::stream()), "Domain::to_device");
}

template <class T>
void Domain<T>::to_device_zero() {
    u_zero = false;
    DeviceArray<T>* arrays[] = {u, requires_duplicate ? uprev : nullptr, f, r, &north->arr, &south->arr,
                                &east->arr, &west->arr, &top->arr, &bottom->arr};
    for (DeviceArray<T>* a : arrays) {
        if (!a) continue;
        a->mark_on_device();
        a->init_zero();
    }
    if (u_alt) {
        const mgp_array d = u->desc();
        mgp_array alt = d;
        alt.data = u_alt;
        mgpx::check(mgpx::Ops<T>::fill(&alt, 0.0, mgpx::stream()), "Domain::to_device_zero");
    }
    mgpx::check(mgp_stream_sync(mgpx::stream()), "Domain::to_device_zero");
}

template <class T>
void Domain<T>::to_host() {
    this->r->to_host();
    this->u->to_host();
    if (requires_duplicate) this->uprev->to_host();
}

template <class T>
void Domain<T>::save(std::string ufilename) {
    this->u->print(settings, ufilename.c_str());
    std::cout << "Saved " << ufilename << std::endl;
}

template <class T>
void Domain<T>::save_all(std::string ufilename, std::string ffilename, std::string rfilename) {
    this->u->print(settings, ufilename.c_str());
    std::cout << "Saved " << ufilename << std::endl;
    this->f->print(settings, ffilename.c_str());
    std::cout << "Saved " << ffilename << std::endl;
    this->r->print(settings, rfilename.c_str());
    std::cout << "Saved " << rfilename << std::endl;
}

template <class T>
void Domain<T>::save_halo() {
    this->u->print_halo(settings, "results/u_halo.vtk");
    this->r->print_halo(settings, "results/r_halo.vtk");
}

template <class T>
void Domain<T>::save_restrict_prolong(Restriction<T>& restriction, Prolongation<T>& prolongation) {
    DeviceArray<T> u_small(settings.dev, settings.dims[0] / 2 + 1, settings.dims[1] / 2 + 1, settings.dims[2] / 2 + 1);
    DeviceArray<T> u_large(settings.dev, settings.dims[0], settings.dims[1], settings.dims[2]);
    u_small.to_device();
    u_large.to_device();
    restriction.restrict_to(*(this->u), u_small);
    prolongation.prolong(u_small, u_large);
    u_small.to_host();
    u_large.to_host();
    u_small.print(settings, "results/u_restricted.vtk");
    u_large.print(settings, "results/u_prolonged.vtk");
}

template <class T>
void Domain<T>::swap_u() {
    if (requires_duplicate) std::swap(this->u, this->uprev);
}

// domain.h:319-326, order N,S,E,W,T,B
template <class T>
void Domain<T>::write_bc_to(DeviceArray<T>& uarr) {
    this->north->write_to(uarr, this->settings);
    this->south->write_to(uarr, this->settings);
    this->east->write_to(uarr, this->settings);
    this->west->write_to(uarr, this->settings);
    this->top->write_to(uarr, this->settings);
    this->bottom->write_to(uarr, this->settings);
}

// domain.h:329-342: f sampled at x0 + i h with x0 and h of type T
template <class T>
void Domain<T>::init_f(funptr ffun) {
    const T x0 = settings.origin[0];
    const T y0 = settings.origin[1];
    const T z0 = settings.origin[2];
    const T h = settings.h;
    for (int_t i = 0; i < this->f->shape[0]; i++)
        for (int_t j = 0; j < this->f->shape[1]; j++)
            for (int_t k = 0; k < this->f->shape[2]; k++)
                this->f->at[this->f->idx(i, j, k)] = ffun(x0 + ((T)i) * h, y0 + ((T)j) * h, z0 + ((T)k) * h);
}

// neumann.h:133-144: refresh the ghosts of u (or uprev for Jacobi)
template <class T>
void Neumann<T>::update(Domain<T>& domain, bool previous, bool fetch_neighbor) {
    (void)fetch_neighbor;
    if (previous)
        this->write_to(*domain.uprev, domain.settings);
    else
        this->write_to(*domain.u, domain.settings);
}
}  // namespace Poisson
#endif
