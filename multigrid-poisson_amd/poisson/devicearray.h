// Array with a GPU copy (reference include/devicearray.h:11-240), on
// hipMalloc'd memory through the mgp C ABI instead of omp_target_alloc.
//
// Semantics kept from the reference: the host mirror `at` always exists;
// init_zero/add/infinity_norm act on the host copy until the first
// to_device() and on the device copy afterwards.  Device work is enqueued on
// the current stream (mgpx::stream()); to_device/to_host/infinity_norm wait
// for it.  Errors are thrown (the reference printed to cerr and continued).
#ifndef DEVICE_ARRAY
#define DEVICE_ARRAY

#include <fstream>
#include <iostream>
#include <stdexcept>

#include "array.h"
#include "halo.h"
#include "mgp_cxx.h"
#include "settings.h"

namespace Poisson {
template <class T>
class DeviceArray : public Array<T> {
private:
    void device_allocator();
    bool on_device = false;

protected:
    uint_t host = -1;

public:
    uint_t device;
    T* devptr;

    DeviceArray(uint_t _device, uint_t i, uint_t j, uint_t k);
    DeviceArray(Settings& settings, Halo& _halo);
    ~DeviceArray();

    void to_device();
    void device_to_device(DeviceArray<T>& devarr);
    void to_host();
    void init_zero();
    void add(const DeviceArray<T>& arr);
    T infinity_norm() const;

    // ---- additions
    mgp_array desc() const;  // C-ABI descriptor of the device copy
    bool is_on_device() const { return on_device; }
    void mark_on_device() { on_device = true; }
};

template <class T>
DeviceArray<T>::DeviceArray(uint_t _device, uint_t i, uint_t j, uint_t k) : Array<T>(i, j, k), device(_device) {
    this->device_allocator();
}

template <class T>
DeviceArray<T>::DeviceArray(Settings& settings, Halo& _halo)
    : Array<T>(settings.dims[0], settings.dims[1], settings.dims[2], _halo), device(settings.dev) {
    this->device_allocator();
}

template <class T>
DeviceArray<T>::~DeviceArray() {
    mgpx::dev_free(this->devptr);
}

template <class T>
void DeviceArray<T>::device_allocator() {
    mgpx::require_device();  // devicearray.h:117-119
    void* p = nullptr;
    mgpx::check(mgpx::dev_alloc(&p, (size_t)this->size * sizeof(T)), "Error allocating DeviceArray");
    this->devptr = static_cast<T*>(p);
}

template <class T>
mgp_array DeviceArray<T>::desc() const {
    mgp_array a;
    a.data = this->devptr;
    for (int d = 0; d < 3; d++) a.shape[d] = this->shape[d];
    a.halo[0] = this->halo.north;
    a.halo[1] = this->halo.south;
    a.halo[2] = this->halo.east;
    a.halo[3] = this->halo.west;
    a.halo[4] = this->halo.top;
    a.halo[5] = this->halo.bottom;
    for (int d = 0; d < 3; d++) a.stride[d] = this->stride[d];
    return a;
}

template <class T>
void DeviceArray<T>::to_device() {
    void* s = mgpx::stream();
    mgpx::check(mgp_memcpy_h2d(this->devptr, this->at, (size_t)this->size * sizeof(T), s), "H2D copy");
    mgpx::check(mgp_stream_sync(s), "H2D copy");
    this->on_device = true;
}

// Whole-array copy to another device array (devicearray.h:78-104; on one GPU
// a device-to-device copy)
template <class T>
void DeviceArray<T>::device_to_device(DeviceArray<T>& devarr) {
    if (devarr.size != this->size) throw std::invalid_argument("device_to_device: size mismatch");
    void* s = mgpx::stream();
    mgpx::check(mgp_memcpy_d2d(devarr.devptr, this->devptr, (size_t)this->size * sizeof(T), s), "D2D copy");
}

template <class T>
void DeviceArray<T>::to_host() {
    void* s = mgpx::stream();
    mgpx::check(mgp_memcpy_d2h(this->at, this->devptr, (size_t)this->size * sizeof(T), s), "D2H copy");
    mgpx::check(mgp_stream_sync(s), "D2H copy");
}

template <class T>
void DeviceArray<T>::init_zero() {
    if (!this->on_device) {
        Array<T>::init_zero();
    } else {
        const mgp_array a = desc();
        mgpx::check(mgpx::Ops<T>::fill(&a, 0.0, mgpx::stream()), "init_zero");
    }
}

template <class T>
void DeviceArray<T>::add(const DeviceArray<T>& arr) {
    if (!this->on_device) {
        Array<T>::add(arr);
    } else {
        const mgp_array a = desc(), b = arr.desc();
        mgpx::check(mgpx::Ops<T>::add(&a, &b, mgpx::stream()), "add");
    }
}

namespace detail {
// A device double and its pinned host mirror for scalar results.
struct Scalar {
    double* dev = nullptr;
    double* host = nullptr;
    // direct: never from an arena (a thread's scalar lives until the thread
    // ends, which may be after an arena installed on the thread is gone)
    explicit Scalar(bool direct = false) {
        mgpx::ArenaScope scope(direct ? nullptr : mgpx::device_arena_slot(),
                               direct ? nullptr : mgpx::host_arena_slot());
        void* p = nullptr;
        mgpx::check(mgpx::dev_alloc(&p, 8 * sizeof(double)), "scalar alloc");
        dev = static_cast<double*>(p);
        mgpx::check(mgpx::host_alloc(&p, 8 * sizeof(double)), "scalar alloc");
        host = static_cast<double*>(p);
    }
    ~Scalar() {
        mgpx::dev_free(dev);
        mgpx::host_free(host);
    }
    Scalar(const Scalar&) = delete;
    Scalar& operator=(const Scalar&) = delete;
};
inline Scalar& thread_scalar() {
    thread_local Scalar s(true);
    return s;
}
}  // namespace detail

template <class T>
T DeviceArray<T>::infinity_norm() const {
    if (!this->on_device) {
        std::cout << "called Array<T>::infinity_norm()" << std::endl;  // devicearray.h:200-201
        return Array<T>::infinity_norm();
    }
    void* s = mgpx::stream();
    detail::Scalar& sc = detail::thread_scalar();
    const mgp_array a = desc();
    mgpx::check(mgp_zero_scalars(sc.dev, 1, s), "infinity_norm");
    mgpx::check(mgpx::Ops<T>::norm_inf(&a, sc.dev, s), "infinity_norm");
    mgpx::check(mgp_memcpy_d2h(sc.host, sc.dev, sizeof(double), s), "infinity_norm");
    mgpx::check(mgp_stream_sync(s), "infinity_norm");
    return (T)sc.host[0];
}
}  // namespace Poisson
#endif
