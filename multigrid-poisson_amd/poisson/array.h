// Halo-padded host array (reference include/array.h:20-272): C order with x
// slowest and z contiguous, per-face halos, idx() hiding the halo offsets,
// host init_zero/add/infinity_norm and the big-endian VTK writer.
#ifndef HOST_ARRAY
#define HOST_ARRAY

#include <inttypes.h>
#include <stdio.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <new>
#include <string>

#include "definitions.h"
#include "halo.h"
#include "settings.h"

namespace Poisson {

__attribute__((always_inline)) constexpr uint_t idx(const uint_t i, const uint_t j, const uint_t k, const Halo& _halo,
                                                    const uint_t (&_stride)[3]) {
    return (i + _halo.west) * _stride[0] + (j + _halo.south) * _stride[1] + (k + _halo.bottom) * _stride[2];
}

__attribute__((always_inline)) constexpr uint_t idx_halo(const uint_t i, const uint_t j, const uint_t k,
                                                         const uint_t (&_stride)[3]) {
    return i * _stride[0] + j * _stride[1] + k * _stride[2];
}

template <class T>
class Array {
private:
    void allocator();

public:
    T* at;
    const uint_t ndims = 3;
    const uint_t size;
    const uint_t shape[3];
    const uint_t stride[3];
    const Halo halo;

    Array(uint_t i, uint_t j, uint_t k);
    Array(uint_t i, uint_t j, uint_t k, Halo& _halo);
    Array(const Array&) = delete;
    Array& operator=(const Array&) = delete;
    virtual ~Array();

    __attribute__((always_inline)) constexpr uint_t idx(const uint_t i, const uint_t j, const uint_t k) const {
        return Poisson::idx(i, j, k, this->halo, this->stride);
    }
    __attribute__((always_inline)) constexpr uint_t idx_halo(const uint_t i, const uint_t j, const uint_t k) const {
        return Poisson::idx_halo(i, j, k, this->stride);
    }

    void init_zero();
    void add(const Array<T>& arr);
    T infinity_norm() const;
    void print(Settings& settings, const char* str) const;
    void print_halo(Settings& settings, const char* str) const;
};

template <class T>
Array<T>::Array(uint_t i, uint_t j, uint_t k) : size(i * j * k), shape{i, j, k}, stride{j * k, k, 1}, halo(0, 0, 0, 0, 0, 0) {
    this->allocator();
}

template <class T>
Array<T>::Array(uint_t i, uint_t j, uint_t k, Halo& _halo)
    : size((i + _halo.east + _halo.west) * (j + _halo.north + _halo.south) * (k + _halo.top + _halo.bottom)),
      shape{i, j, k},
      stride{(j + _halo.north + _halo.south) * (k + _halo.top + _halo.bottom), (k + _halo.top + _halo.bottom), 1},
      halo(_halo) {
    this->allocator();
}

template <class T>
Array<T>::~Array() {
    std::free(this->at);
}

// Zero-initialised (the reference leaves new T[] uninitialised; a zeroed mirror
// makes every array well defined before its first write).  calloc: large
// mirrors come from zero pages the OS maps on first touch, so a mirror that is
// never used costs no host time.
template <class T>
void Array<T>::allocator() {
    this->at = static_cast<T*>(std::calloc(size > 0 ? size : 1, sizeof(T)));
    if (!this->at) throw std::bad_alloc();
}

template <class T>
void Array<T>::init_zero() {
    for (uint_t i = 0; i < this->size; i++) this->at[i] = 0.0;
}

template <class T>
void Array<T>::add(const Array<T>& arr) {
    if (!(arr.size == this->size)) throw std::invalid_argument("Array size does not match in Array<T>.add()");
    for (uint_t i = 0; i < this->shape[0]; i++)
        for (uint_t j = 0; j < this->shape[1]; j++)
            for (uint_t k = 0; k < this->shape[2]; k++) this->at[this->idx(i, j, k)] += arr.at[arr.idx(i, j, k)];
}

template <class T>
T Array<T>::infinity_norm() const {
    T res = 0.0;
    for (uint_t i = 0; i < this->shape[0]; i++)
        for (uint_t j = 0; j < this->shape[1]; j++)
            for (uint_t k = 0; k < this->shape[2]; k++) {
                T abselem = std::abs(this->at[this->idx(i, j, k)]);
                res = std::max(res, abselem);
            }
    return res;
}

inline int is_little_endian(void) {
    int num = 1;
    return (*((char*)&num) == 1);
}

namespace detail {
// Writes one element as a big-endian IEEE double, as VTK "double" data.  The
// reference reinterprets the element's own bytes as 8 bytes (array.h:242),
// which is only right for T = double; this writer converts to double first.
template <class T>
inline size_t write_be_double(FILE* f, T v) {
    double d = (double)v;
    uint64_t crnt;
    std::memcpy(&crnt, &d, 8);
    if (is_little_endian()) {
        crnt = (crnt & 0x00000000FFFFFFFFull) << 32 | (crnt & 0xFFFFFFFF00000000ull) >> 32;
        crnt = (crnt & 0x0000FFFF0000FFFFull) << 16 | (crnt & 0xFFFF0000FFFF0000ull) >> 16;
        crnt = (crnt & 0x00FF00FF00FF00FFull) << 8 | (crnt & 0xFF00FF00FF00FF00ull) >> 8;
    }
    return fwrite(&crnt, sizeof(uint64_t), 1, f);
}

inline void vtk_header(FILE* f, long nx, long ny, long nz, double ox, double oy, double oz, double h, unsigned long n) {
    fprintf(f, "# vtk DataFile Version 3.0\n");
    fprintf(f, "saved from function print_vtk.\n");
    fprintf(f, "BINARY\n");
    fprintf(f, "DATASET STRUCTURED_POINTS\n");
    fprintf(f, "DIMENSIONS %d %d %d\n", (int)nx, (int)ny, (int)nz);
    fprintf(f, "ORIGIN %f %f %f\n", ox, oy, oz);
    fprintf(f, "SPACING %f %f %f\n", h, h, h);
    fprintf(f, "POINT_DATA %lu\n", n);
    fprintf(f, "SCALARS %s %s 1\n", "gray", "double");
    fprintf(f, "LOOKUP_TABLE default\n");
}
}  // namespace detail

// VTK STRUCTURED_POINTS, binary big-endian doubles, x fastest (array.h:148-203)
template <class T>
void Array<T>::print(Settings& settings, const char* fname) const {
    FILE* f_ptr = fopen(fname, "w");
    if (f_ptr == NULL) {
        perror("No output! fopen()");
        return;
    }
    const uint_t items = this->shape[0] * this->shape[1] * this->shape[2];
    detail::vtk_header(f_ptr, this->shape[0], this->shape[1], this->shape[2], settings.origin[0], settings.origin[1],
                       settings.origin[2], settings.h, (unsigned long)items);
    uint_t written = 0;
    for (uint_t k = 0; k < this->shape[2]; ++k)
        for (uint_t j = 0; j < this->shape[1]; ++j)
            for (uint_t i = 0; i < this->shape[0]; ++i) written += detail::write_be_double(f_ptr, this->at[this->idx(i, j, k)]);
    if (written != items) std::cerr << "Failed to print file " << fname << std::endl;
    fclose(f_ptr);
}

template <class T>
void Array<T>::print_halo(Settings& settings, const char* fname) const {
    FILE* f_ptr = fopen(fname, "w");
    if (f_ptr == NULL) {
        perror("No output! fopen()");
        return;
    }
    const uint_t nx = this->shape[0] + this->halo.east + this->halo.west;
    const uint_t ny = this->shape[1] + this->halo.north + this->halo.south;
    const uint_t nz = this->shape[2] + this->halo.top + this->halo.bottom;
    std::cout << "Printing file " << fname << " with shape (" << nx << "," << ny << "," << nz << ")" << std::endl;
    detail::vtk_header(f_ptr, nx, ny, nz, settings.origin[0] - ((double)this->halo.west) * settings.h,
                       settings.origin[1] - ((double)this->halo.south) * settings.h,
                       settings.origin[2] - ((double)this->halo.bottom) * settings.h, settings.h,
                       (unsigned long)this->size);
    uint_t written = 0;
    for (uint_t k = 0; k < nz; ++k)
        for (uint_t j = 0; j < ny; ++j)
            for (uint_t i = 0; i < nx; ++i) written += detail::write_be_double(f_ptr, this->at[this->idx_halo(i, j, k)]);
    if (written != this->size) std::cerr << "Failed to print file " << fname << std::endl;
    fclose(f_ptr);
}
}  // namespace Poisson
#endif
