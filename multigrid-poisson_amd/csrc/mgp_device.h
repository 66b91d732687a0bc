// Device-side views and helpers shared by the mgp kernels (gfx950).
//
// Arrays are the reference's halo-padded C-order layout (x slowest, z
// contiguous): element (i,j,k) of the shape at base + i*s0 + j*s1 + k, where
// base folds in the WEST/SOUTH/BOTTOM halo offsets.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mgp.h"

namespace mgp {

template <class T>
struct View {
    T* p;        // points at element (0,0,0) of the shape (halo offsets applied)
    int64_t n0, n1, n2;
    int64_t s0, s1;
    __device__ __forceinline__ int64_t ix(int64_t i, int64_t j, int64_t k) const { return i * s0 + j * s1 + k; }
    __device__ __forceinline__ T& at(int64_t i, int64_t j, int64_t k) const { return p[ix(i, j, k)]; }
};

template <class T>
inline View<T> view_of(const mgp_array& a) {
    View<T> v;
    const int64_t base = a.halo[MGP_WEST] * a.stride[0] + a.halo[MGP_SOUTH] * a.stride[1] + a.halo[MGP_BOTTOM];
    v.p = static_cast<T*>(a.data) + base;
    v.n0 = a.shape[0];
    v.n1 = a.shape[1];
    v.n2 = a.shape[2];
    v.s0 = a.stride[0];
    v.s1 = a.stride[1];
    return v;
}

// Boundary faces as seen by a kernel that forms Neumann ghosts itself.
template <class T>
struct Faces {
    const T* g[6];
    int32_t neumann[6];
    T two_h;
};

template <class T>
inline Faces<T> faces_of(const mgp_faces& f) {
    Faces<T> F;
    for (int q = 0; q < 6; q++) {
        F.g[q] = static_cast<const T*>(f.g[q]);
        F.neumann[q] = f.type[q] == MGP_NEUMANN;
    }
    F.two_h = (T)(2.0 * f.h);  // include/neumann.h:108
    return F;
}

// Value of the stencil neighbour of in-shape point (i,j,k) one step along
// `axis` in direction `dir`; when that neighbour lies outside the shape it is
// the Neumann ghost sign*2h*g + u(mirror) (include/neumann.h:112-130), formed
// from the current u.  Only called for faces whose points are updated, i.e.
// Neumann faces (Dirichlet boundary nodes are never updated, so their
// neighbours beyond the face are never needed).
template <class T>
__device__ __forceinline__ T neighbour(const View<T>& u, const Faces<T>& F, int64_t i, int64_t j, int64_t k, int axis,
                                       int dir) {
    int64_t a[3] = {i, j, k};
    const int64_t n[3] = {u.n0, u.n1, u.n2};
    a[axis] += dir;
    if (a[axis] >= 0 && a[axis] < n[axis]) return u.at(a[0], a[1], a[2]);
    // ghost: face index and mirror point
    int face;
    T sign;
    if (axis == 0) face = dir > 0 ? MGP_EAST : MGP_WEST;
    else if (axis == 1) face = dir > 0 ? MGP_NORTH : MGP_SOUTH;
    else face = dir > 0 ? MGP_TOP : MGP_BOTTOM;
    sign = dir > 0 ? (T)1.0 : (T)-1.0;
    a[axis] -= 2 * dir;  // mirror: 1 for the low face, n-2 for the high face
    int64_t gi;
    if (axis == 0) gi = j * u.n2 + k;          // (1, ny, nz)
    else if (axis == 1) gi = i * u.n2 + k;     // (nx, 1, nz)
    else gi = i * u.n1 + j;                    // (nx, ny, 1)
    const T g = F.g[face] ? F.g[face][gi] : (T)0.0;
    return sign * F.two_h * g + u.at(a[0], a[1], a[2]);
}

}  // namespace mgp
