// Per-point arithmetic of the multigrid operators, shared by the per-level
// kernels (mgp_kernels.hip) and the batched kernels (mgp_batch.hip) so that
// both produce the same bits for the same inputs.  Reference-mode expressions
// keep the reference's operand order and C++ types; the files including this
// are built with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mgp.h"
#include "mgp_device.h"

namespace mgp {

template <class T>
struct RelaxConsts {
    T one_omega, omega_sixth, hsq;
    RelaxConsts(double omega_in, double h) {
        const T omega = (T)omega_in;
        one_omega = (T)(1.0 - omega);
        omega_sixth = (T)(omega / 6.0);
        hsq = (T)(h * h);
    }
};

// One smoother update of point c from the current values around it.
// Reference: include/gaussseidel.h:58-60 / jacobi.h:66-69
//   (1-w) v + w/6 (vW + vE + vS + vN + vB + vT - h^2 f)
// Consistent: v + w/6 * fma(-h^2, f, sum(v_nb - v))
template <class T, int MODE>
__device__ __forceinline__ T relax_point(T vc, T vW, T vE, T vS, T vN, T vB, T vT, T fc, T one_omega, T omega_sixth,
                                         T hsq) {
    if (MODE == MGP_MODE_REFERENCE) {
        return one_omega * vc + omega_sixth * (vW + vE + vS + vN + vB + vT - hsq * fc);
    } else {
        const T s = ((vW - vc) + (vE - vc)) + ((vS - vc) + (vN - vc)) + ((vB - vc) + (vT - vc));
        return fma(omega_sixth, fma(-hsq, fc, s), vc);
    }
}

// include/residual.h:58-79.  Reference: r = f - (sum_nb - 6.0*u)/hsq with the
// literal 6.0 a double, as in the reference.  Consistent:
// r = fma(-1/h^2, sum(u_nb - u), f).
template <class T, int MODE>
__device__ __forceinline__ T residual_point(const View<T>& u, int64_t c, T fc, T hsq, T inv_h2) {
    const T* p = u.p;
    if (MODE == MGP_MODE_REFERENCE) {
        return fc - (p[c - u.s0] + p[c + u.s0] + p[c - u.s1] + p[c + u.s1] + p[c - 1] + p[c + 1] - 6.0 * p[c]) / hsq;
    } else {
        const T uc = p[c];
        const T s = ((p[c - u.s0] - uc) + (p[c + u.s0] - uc)) + ((p[c - u.s1] - uc) + (p[c + u.s1] - uc)) +
                    ((p[c - 1] - uc) + (p[c + 1] - uc));
        return fma(-inv_h2, s, fc);
    }
}

// ---- iterate held as hi + lo (FP64 hi, FP32 lo) in mixed-precision
// refinement: a solution carried to about 2^-77 relative, so that the FP64
// residual of it is not limited by the rounding of u to FP64 (no counterpart
// in the reference, whose u is one array of T).

// Error-free sum (Knuth's TwoSum): s = fl(a + b) and a + b = s + e exactly
// (needs separate roundings: the including files build with -ffp-contract=off)
__device__ __forceinline__ double two_sum(double a, double b, double& e) {
    const double s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
    return s;
}

// The refinement update u += e of u = hi + lo: hi' = fl(hi + t) with
// t = fl(e + lo), lo' = (float) of the exact rounding error of hi'
__device__ __forceinline__ void hilo_update(double& hi, float& lo, float e) {
    double err;
    hi = two_sum(hi, (double)e + (double)lo, err);
    lo = (float)err;
}

// Residual of u = hi + lo at one point, consistent-mode order:
// r = fma(-1/h^2, S(hi) + S(lo), f), S(v) = ((vW-vc)+(vE-vc)) + ((vS-vc)+(vN-vc))
// + ((vB-vc)+(vT-vc)); the lo differences are taken in FP64
__device__ __forceinline__ double hilo_residual_point(double hc, double hW, double hE, double hS, double hN, double hB,
                                                      double hT, float lc, float lW, float lE, float lS, float lN,
                                                      float lB, float lT, double fc, double inv_h2) {
    const double sh = ((hW - hc) + (hE - hc)) + ((hS - hc) + (hN - hc)) + ((hB - hc) + (hT - hc));
    const double c = (double)lc;
    const double sl = (((double)lW - c) + ((double)lE - c)) + (((double)lS - c) + ((double)lN - c)) +
                      (((double)lB - c) + ((double)lT - c));
    return fma(-inv_h2, sh + sl, fc);
}

// Neumann ghost of hi + lo: (p + mirror) with p = sign*2h*g, the rounding of
// p + hi(mirror) carried into lo
__device__ __forceinline__ void hilo_ghost(double p, double hm, float lm, double& hg, float& lg) {
    double err;
    hg = two_sum(p, hm, err);
    lg = (float)((double)lm + err);
}

// The grid transfers below take any view type V with at(i,j,k) and n0..n2
// (View<T>, or FP16 storage read as float).
//
// Full weighting at coarse point (i,j,k): (1,2,1)^3/64 of the fine values
// with neighbour indices clamped to [0, n-1]; the three z-planes summed in the
// reference's term order (T tmp = plane; tmp += plane),
// include/fullweighting.h:59-121
template <class T, class V = View<T>>
__device__ __forceinline__ T fw_point(const V& in, int64_t i, int64_t j, int64_t k) {
    const int64_t n0 = in.n0, n1 = in.n1, n2 = in.n2;
    const int64_t ii = min(n0 - 1, 2 * i), jj = min(n1 - 1, 2 * j), kk = min(n2 - 1, 2 * k);
    const int64_t im = max((int64_t)0, ii - 1), ip = min(n0 - 1, ii + 1);
    const int64_t jm = max((int64_t)0, jj - 1), jp = min(n1 - 1, jj + 1);
    const int64_t km = max((int64_t)0, kk - 1), kp = min(n2 - 1, kk + 1);
#define A(a, b, c) in.at(a, b, c)
    T tmp = 1.0 * A(im, jm, km) + 2.0 * A(im, jj, km) + 1.0 * A(im, jp, km) + 2.0 * A(ii, jm, km) +
            4.0 * A(ii, jj, km) + 2.0 * A(ii, jp, km) + 1.0 * A(ip, jm, km) + 2.0 * A(ip, jj, km) +
            1.0 * A(ip, jp, km);
    tmp += 2.0 * A(im, jm, kk) + 4.0 * A(im, jj, kk) + 2.0 * A(im, jp, kk) + 4.0 * A(ii, jm, kk) +
           8.0 * A(ii, jj, kk) + 4.0 * A(ii, jp, kk) + 2.0 * A(ip, jm, kk) + 4.0 * A(ip, jj, kk) +
           2.0 * A(ip, jp, kk);
    tmp += 1.0 * A(im, jm, kp) + 2.0 * A(im, jj, kp) + 1.0 * A(im, jp, kp) + 2.0 * A(ii, jm, kp) +
           4.0 * A(ii, jj, kp) + 2.0 * A(ii, jp, kp) + 1.0 * A(ip, jm, kp) + 2.0 * A(ip, jj, kp) +
           1.0 * A(ip, jp, kp);
#undef A
    return tmp / 64.0;
}

// Trilinear value at fine point (i,j,k), composed exactly as the reference's
// four passes compose it (include/trilinearinterpolation.h:50-133): injection
// at even points, then x-, y- and z-midpoints, each 0.5*(left + right) of the
// previous passes' values.
template <class T>
__device__ __forceinline__ T half_sum(T a, T b) {
    return 0.5 * (a + b);
}

template <class T, class V = View<T>>
__device__ __forceinline__ T tri_xy(const V& c, int64_t a, int64_t b, int64_t cc, int pi, int pj) {
    if (!pi && !pj) return c.at(a, b, cc);
    if (pi && !pj) return half_sum<T>(c.at(a, b, cc), c.at(a + 1, b, cc));
    if (!pi && pj) return half_sum<T>(c.at(a, b, cc), c.at(a, b + 1, cc));
    return half_sum<T>(half_sum<T>(c.at(a, b, cc), c.at(a + 1, b, cc)),
                       half_sum<T>(c.at(a, b + 1, cc), c.at(a + 1, b + 1, cc)));
}

template <class T, class V = View<T>>
__device__ __forceinline__ T trilinear(const V& c, int64_t i, int64_t j, int64_t k) {
    const int pi = (int)(i & 1), pj = (int)(j & 1), pk = (int)(k & 1);
    const int64_t a = i >> 1, b = j >> 1, cc = k >> 1;
    const T lo = tri_xy<T, V>(c, a, b, cc, pi, pj);
    if (!pk) return lo;
    return half_sum<T>(lo, tri_xy<T, V>(c, a, b, cc + 1, pi, pj));
}

}  // namespace mgp
