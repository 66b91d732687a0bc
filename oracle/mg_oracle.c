/*
 * CPU ORACLE — test infrastructure only (see mg_oracle.h for the contract).
 *
 * Every function cites the reference file:line it restates.  Reference-mode
 * arithmetic reproduces the reference's expression order term by term; build
 * with -ffp-contract=off so that no multiply-add is fused behind our back.
 * Consistent-mode arithmetic is this project's own (explicit fma()).
 */
#define _GNU_SOURCE
#include "mg_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define MIN(a, b) (((a) < (b)) ? (a) : (b))
#define MAX(a, b) (((a) > (b)) ? (a) : (b))

/* ------------------------------------------------------------------ layout */

/* include/array.h:135-141: stride = {(ny+hN+hS)(nz+hT+hB), nz+hT+hB, 1} */
void og_layout_make(og_layout* L, int64_t nx, int64_t ny, int64_t nz, const int64_t halo[6]) {
    static const int64_t zero[6] = {0, 0, 0, 0, 0, 0};
    if (!halo) halo = zero;
    L->n[0] = nx;
    L->n[1] = ny;
    L->n[2] = nz;
    for (int q = 0; q < 6; q++) L->halo[q] = halo[q];
    L->s[2] = 1;
    L->s[1] = nz + halo[OG_T] + halo[OG_B];
    L->s[0] = (ny + halo[OG_N] + halo[OG_S]) * L->s[1];
    L->size = (nx + halo[OG_E] + halo[OG_W]) * L->s[0];
}

/* include/array.h:84-88 */
static inline int64_t IX(const og_layout* L, int64_t i, int64_t j, int64_t k) {
    return (i + L->halo[OG_W]) * L->s[0] + (j + L->halo[OG_S]) * L->s[1] + (k + L->halo[OG_B]);
}

/* ---------------------------------------------------------------- problems */
/* Problem 0: u = sin(2 pi x) sin(2 pi y) sin(2 pi z), f = -12 pi^2 u.
 * Problems 1-3 restate src/problem_definition.cpp:1-76 (PROBLEM1..3). */
double og_ufun(int problem, double x, double y, double z) {
    switch (problem) {
    case 0: { const double p = 2.0 * M_PI; return sin(p * x) * sin(p * y) * sin(p * z); }
    case 1: return sin(3.0 * x) * sin(5.3 * y) * sin(-7.8 * z);
    case 2: return x * x * x * y * y * z;
    default: return cos(x * z * z) * sin(y * y * y);
    }
}
double og_ffun(int problem, double x, double y, double z) {
    switch (problem) {
    case 0: { const double p = 2.0 * M_PI; return -3.0 * p * p * sin(p * x) * sin(p * y) * sin(p * z); }
    case 1: return -(3.0 * 3.0 + 5.3 * 5.3 + -7.8 * -7.8) * sin(3.0 * x) * sin(5.3 * y) * sin(-7.8 * z);
    case 2: return 2.0 * x * x * x * z + 6.0 * x * y * y * z;
    default: {
        /* the reference's GCC build evaluates this f with two sincos() calls
         * (GCC merges sin/cos pairs of one argument); spelled out so the
         * values do not depend on that optimisation */
        double sy, cy, sxz, cxz;
        sincos(y * y * y, &sy, &cy);
        sincos(x * z * z, &sxz, &cxz);
        return ((-4.0 * x * x * z * z - 9.0 * y * y * y * y - z * z * z * z) * sy + 6.0 * y * cy) * cxz -
               2.0 * x * sxz * sy;
    }
    }
}
double og_dudx(int problem, double x, double y, double z) {
    switch (problem) {
    case 0: { const double p = 2.0 * M_PI; return p * cos(p * x) * sin(p * y) * sin(p * z); }
    case 1: return 3.0 * cos(3.0 * x) * sin(5.3 * y) * sin(-7.8 * z);
    case 2: return 3.0 * x * x * y * y * z;
    default: return -z * z * sin(x * z * z) * sin(y * y * y);
    }
}
double og_dudy(int problem, double x, double y, double z) {
    switch (problem) {
    case 0: { const double p = 2.0 * M_PI; return p * sin(p * x) * cos(p * y) * sin(p * z); }
    case 1: return 5.3 * sin(3.0 * x) * cos(5.3 * y) * sin(-7.8 * z);
    case 2: return 2.0 * x * x * x * y * z;
    default: return 3.0 * cos(x * z * z) * y * y * cos(y * y * y);
    }
}
double og_dudz(int problem, double x, double y, double z) {
    switch (problem) {
    case 0: { const double p = 2.0 * M_PI; return p * sin(p * x) * sin(p * y) * cos(p * z); }
    case 1: return -7.8 * sin(3.0 * x) * sin(5.3 * y) * cos(-7.8 * z);
    case 2: return x * x * x * y * y;
    default: return -2 * x * z * sin(x * z * z) * sin(y * y * y);
    }
}

/* --------------------------------------------------------------- operators */

/* Reference: include/residual.h:58-79
 *   r = f - (uW + uE + uS + uN + uB + uT - 6u) / h^2   (left to right)
 * Consistent: r = fma(-(1/h^2), ((dW+dE)+(dS+dN))+(dB+dT), f), d = u_nb - u_c. */
void og_residual(int mode, const og_layout* L, const double* u, const double* f, double* r,
                 const int64_t box[6], double h) {
    const double hsq = h * h;
    const double inv_h2 = 1.0 / hsq;
    const int64_t s0 = L->s[0], s1 = L->s[1];
#pragma omp parallel for schedule(static)
    for (int64_t i = box[0]; i < box[1]; i++)
        for (int64_t j = box[2]; j < box[3]; j++)
            for (int64_t k = box[4]; k < box[5]; k++) {
                const int64_t c = IX(L, i, j, k);
                if (mode == OG_MODE_REFERENCE) {
                    r[c] = f[c] - (u[c - s0] + u[c + s0] + u[c - s1] + u[c + s1] + u[c - 1] + u[c + 1] -
                                   6.0 * u[c]) / hsq;
                } else {
                    const double uc = u[c];
                    const double s = ((u[c - s0] - uc) + (u[c + s0] - uc)) + ((u[c - s1] - uc) + (u[c + s1] - uc)) +
                                     ((u[c - 1] - uc) + (u[c + 1] - uc));
                    r[c] = fma(-inv_h2, s, f[c]);
                }
            }
}

/* Colour of a point.  Reference (include/gaussseidel.h:56,81): parity of the
 * halo-inclusive linear index, red = even.  Consistent: parity of i+j+k. */
static inline int colour_of(int mode, const og_layout* L, int64_t i, int64_t j, int64_t k) {
    if (mode == OG_MODE_REFERENCE) return (int)(IX(L, i, j, k) % 2);
    return (int)((i + j + k) & 1);
}

/* True when a colour class is a proper checkerboard (no same-colour
 * neighbours), so the in-place half-sweep is order independent. */
static int colouring_is_checkerboard(int mode, const og_layout* L) {
    if (mode != OG_MODE_REFERENCE) return 1;
    return (L->s[0] & 1) && (L->s[1] & 1);
}

static inline void gs_point(int mode, const og_layout* L, double* u, const double* f, int64_t c,
                            double one_omega, double omega_sixth, double hsq) {
    const int64_t s0 = L->s[0], s1 = L->s[1];
    if (mode == OG_MODE_REFERENCE) {
        u[c] = one_omega * u[c] + omega_sixth * (u[c - s0] + u[c + s0] + u[c - s1] + u[c + s1] + u[c - 1] +
                                                 u[c + 1] - hsq * f[c]);
    } else {
        const double uc = u[c];
        const double s = ((u[c - s0] - uc) + (u[c + s0] - uc)) + ((u[c - s1] - uc) + (u[c + s1] - uc)) +
                         ((u[c - 1] - uc) + (u[c + 1] - uc));
        u[c] = fma(omega_sixth, fma(-hsq, f[c], s), uc);
    }
}

/* Reference: include/gaussseidel.h:28-94 — one half-sweep (colour 0 = red)
 * over the box, in place, u = (1-w)u + w/6 (sum_nb - h^2 f).
 * A non-checkerboard colouring (even strides) is a data race in the
 * reference; it is run here in the reference's sequential (one-thread) order. */
void og_gs_half(int mode, const og_layout* L, double* u, const double* f, const int64_t box[6],
                double omega, double h, int colour) {
    const double one_omega = 1.0 - omega;
    const double omega_sixth = omega / 6.0;
    const double hsq = h * h;
    if (colouring_is_checkerboard(mode, L)) {
#pragma omp parallel for schedule(static)
        for (int64_t i = box[0]; i < box[1]; i++)
            for (int64_t j = box[2]; j < box[3]; j++)
                for (int64_t k = box[4]; k < box[5]; k++)
                    if (colour_of(mode, L, i, j, k) == colour)
                        gs_point(mode, L, u, f, IX(L, i, j, k), one_omega, omega_sixth, hsq);
    } else {
        for (int64_t i = box[0]; i < box[1]; i++)
            for (int64_t j = box[2]; j < box[3]; j++)
                for (int64_t k = box[4]; k < box[5]; k++)
                    if (colour_of(mode, L, i, j, k) == colour)
                        gs_point(mode, L, u, f, IX(L, i, j, k), one_omega, omega_sixth, hsq);
    }
}

/* Reference: include/jacobi.h:59-81 — u = (1-w) v; u += w/6 (sum_nb(v) - h^2 f) */
void og_jacobi(int mode, const og_layout* L, double* u, const double* v, const double* f,
               const int64_t box[6], double omega, double h) {
    const double one_omega = 1.0 - omega;
    const double omega_sixth = omega / 6.0;
    const double hsq = h * h;
    const int64_t s0 = L->s[0], s1 = L->s[1];
#pragma omp parallel for schedule(static)
    for (int64_t i = box[0]; i < box[1]; i++)
        for (int64_t j = box[2]; j < box[3]; j++)
            for (int64_t k = box[4]; k < box[5]; k++) {
                const int64_t c = IX(L, i, j, k);
                if (mode == OG_MODE_REFERENCE) {
                    double t = one_omega * v[c];
                    t += omega_sixth * (v[c - s0] + v[c + s0] + v[c - s1] + v[c + s1] + v[c - 1] + v[c + 1] -
                                        hsq * f[c]);
                    u[c] = t;
                } else {
                    const double vc = v[c];
                    const double s = ((v[c - s0] - vc) + (v[c + s0] - vc)) + ((v[c - s1] - vc) + (v[c + s1] - vc)) +
                                     ((v[c - 1] - vc) + (v[c + 1] - vc));
                    u[c] = fma(omega_sixth, fma(-hsq, f[c], s), vc);
                }
            }
}

/* Reference: include/injection.h:47-66 — out(i,j,k) = in(2i,2j,2k) over out's shape */
void og_restrict_inject(const og_layout* Li, const double* in, const og_layout* Lo, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Lo->n[0]; i++)
        for (int64_t j = 0; j < Lo->n[1]; j++)
            for (int64_t k = 0; k < Lo->n[2]; k++) out[IX(Lo, i, j, k)] = in[IX(Li, 2 * i, 2 * j, 2 * k)];
}

/* Reference: include/fullweighting.h:59-121 — (1,2,1)^3/64 with neighbour
 * indices clamped to [0, n-1]; three z-planes summed in the reference's
 * term order, tmp += plane. */
void og_restrict_fw(const og_layout* Li, const double* in, const og_layout* Lo, double* out) {
    const int64_t n0 = Li->n[0], n1 = Li->n[1], n2 = Li->n[2];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Lo->n[0]; i++)
        for (int64_t j = 0; j < Lo->n[1]; j++)
            for (int64_t k = 0; k < Lo->n[2]; k++) {
                const int64_t ii = MIN(n0 - 1, 2 * i), jj = MIN(n1 - 1, 2 * j), kk = MIN(n2 - 1, 2 * k);
                const int64_t im = MAX(0, ii - 1), ip = MIN(n0 - 1, ii + 1);
                const int64_t jm = MAX(0, jj - 1), jp = MIN(n1 - 1, jj + 1);
                const int64_t km = MAX(0, kk - 1), kp = MIN(n2 - 1, kk + 1);
#define A(a, b, c) in[IX(Li, a, b, c)]
                double tmp = 1.0 * A(im, jm, km) + 2.0 * A(im, jj, km) + 1.0 * A(im, jp, km) + 2.0 * A(ii, jm, km) +
                             4.0 * A(ii, jj, km) + 2.0 * A(ii, jp, km) + 1.0 * A(ip, jm, km) + 2.0 * A(ip, jj, km) +
                             1.0 * A(ip, jp, km);
                tmp += 2.0 * A(im, jm, kk) + 4.0 * A(im, jj, kk) + 2.0 * A(im, jp, kk) + 4.0 * A(ii, jm, kk) +
                       8.0 * A(ii, jj, kk) + 4.0 * A(ii, jp, kk) + 2.0 * A(ip, jm, kk) + 4.0 * A(ip, jj, kk) +
                       2.0 * A(ip, jp, kk);
                tmp += 1.0 * A(im, jm, kp) + 2.0 * A(im, jj, kp) + 1.0 * A(im, jp, kp) + 2.0 * A(ii, jm, kp) +
                       4.0 * A(ii, jj, kp) + 2.0 * A(ii, jp, kp) + 1.0 * A(ip, jm, kp) + 2.0 * A(ip, jj, kp) +
                       1.0 * A(ip, jp, kp);
#undef A
                out[IX(Lo, i, j, k)] = tmp / 64.0;
            }
}

/* Reference: include/trilinearinterpolation.h:29-134 — inject the coarse
 * values at even points, then midpoints along x, then y, then z, each pass
 * reading the previous passes' results: out = 0.5*(left + right). */
void og_prolong_trilinear(const og_layout* Li, const double* in, const og_layout* Lo, double* out) {
    const int64_t n0 = Lo->n[0], n1 = Lo->n[1], n2 = Lo->n[2];
    const int64_t s0 = Lo->s[0], s1 = Lo->s[1];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Li->n[0]; i++)
        for (int64_t j = 0; j < Li->n[1]; j++)
            for (int64_t k = 0; k < Li->n[2]; k++) out[IX(Lo, 2 * i, 2 * j, 2 * k)] = in[IX(Li, i, j, k)];
#pragma omp parallel for schedule(static)
    for (int64_t i = 1; i < n0 - 1; i += 2)
        for (int64_t j = 0; j < n1; j += 2)
            for (int64_t k = 0; k < n2; k += 2) {
                const int64_t c = IX(Lo, i, j, k);
                out[c] = 0.5 * (out[c - s0] + out[c + s0]);
            }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n0; i++)
        for (int64_t j = 1; j < n1 - 1; j += 2)
            for (int64_t k = 0; k < n2; k += 2) {
                const int64_t c = IX(Lo, i, j, k);
                out[c] = 0.5 * (out[c - s1] + out[c + s1]);
            }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n0; i++)
        for (int64_t j = 0; j < n1; j++)
            for (int64_t k = 1; k < n2 - 1; k += 2) {
                const int64_t c = IX(Lo, i, j, k);
                out[c] = 0.5 * (out[c - 1] + out[c + 1]);
            }
}

/* face array shape per location: include/domain.h:125-190 */
static void face_shape(int loc, const int64_t n[3], int64_t fs[3]) {
    fs[0] = n[0];
    fs[1] = n[1];
    fs[2] = n[2];
    if (loc == OG_E || loc == OG_W) fs[0] = 1;
    if (loc == OG_N || loc == OG_S) fs[1] = 1;
    if (loc == OG_T || loc == OG_B) fs[2] = 1;
}

/* Reference: include/dirichlet.h:50-93 — boundary nodes (index 0 or n-1) = g */
void og_bc_dirichlet(const og_layout* L, double* u, int loc, const double* g) {
    int64_t fs[3], off[3] = {0, 0, 0};
    face_shape(loc, L->n, fs);
    if (loc == OG_T) off[2] = L->n[2] - 1;
    if (loc == OG_N) off[1] = L->n[1] - 1;
    if (loc == OG_E) off[0] = L->n[0] - 1;
    og_layout G;
    og_layout_make(&G, fs[0], fs[1], fs[2], NULL);
    for (int64_t i = 0; i < fs[0]; i++)
        for (int64_t j = 0; j < fs[1]; j++)
            for (int64_t k = 0; k < fs[2]; k++) u[IX(L, i + off[0], j + off[1], k + off[2])] = g[IX(&G, i, j, k)];
}

/* Reference: include/neumann.h:63-131 — ghost = sign*2h*g + u(mirror),
 * sign = -1 on WEST/SOUTH/BOTTOM */
void og_bc_neumann(const og_layout* L, double* u, int loc, const double* g, double h) {
    int64_t fs[3], off[3] = {0, 0, 0}, d[3] = {0, 0, 0};
    double sign = 1.0;
    face_shape(loc, L->n, fs);
    switch (loc) {
    case OG_E: off[0] = L->n[0]; d[0] = -2; break;
    case OG_W: off[0] = -1; d[0] = 2; sign = -1.0; break;
    case OG_N: off[1] = L->n[1]; d[1] = -2; break;
    case OG_S: off[1] = -1; d[1] = 2; sign = -1.0; break;
    case OG_T: off[2] = L->n[2]; d[2] = -2; break;
    default: off[2] = -1; d[2] = 2; sign = -1.0; break;
    }
    const double two_h = 2.0 * h;
    og_layout G;
    og_layout_make(&G, fs[0], fs[1], fs[2], NULL);
    for (int64_t i = 0; i < fs[0]; i++)
        for (int64_t j = 0; j < fs[1]; j++)
            for (int64_t k = 0; k < fs[2]; k++) {
                const int64_t a = i + off[0], b = j + off[1], c = k + off[2];
                u[IX(L, a, b, c)] = sign * two_h * g[IX(&G, i, j, k)] + u[IX(L, a + d[0], b + d[1], c + d[2])];
            }
}

/* Reference: include/neumann.h:146-252 — g_c = R(g_f), then
 * g_c -= (c1 u_f(b) + c2 u_f(b+-1) + c3 u_f(b+-2)) / h_f,
 * c = (-3/2, 2, -1/2) on WEST/SOUTH/BOTTOM and (3/2, -2, 1/2) elsewhere. */
void og_bc_restrict_neumann(int restriction, const og_layout* Lf, const double* u_fine, int loc,
                            const double* g_fine, double* g_coarse, double h_fine) {
    int64_t ffs[3], nc[3], cfs[3];
    face_shape(loc, Lf->n, ffs);
    for (int a = 0; a < 3; a++) nc[a] = Lf->n[a] / 2 + 1;
    face_shape(loc, nc, cfs);
    og_layout GF, GC;
    og_layout_make(&GF, ffs[0], ffs[1], ffs[2], NULL);
    og_layout_make(&GC, cfs[0], cfs[1], cfs[2], NULL);
    if (restriction == OG_FULL_WEIGHTING)
        og_restrict_fw(&GF, g_fine, &GC, g_coarse);
    else
        og_restrict_inject(&GF, g_fine, &GC, g_coarse);
    int64_t off[3] = {0, 0, 0}, d1[3] = {0, 0, 0}, d2[3] = {0, 0, 0};
    switch (loc) {
    case OG_E: off[0] = Lf->n[0] / 2; d1[0] = -1; d2[0] = -2; break;
    case OG_W: d1[0] = 1; d2[0] = 2; break;
    case OG_N: off[1] = Lf->n[1] / 2; d1[1] = -1; d2[1] = -2; break;
    case OG_S: d1[1] = 1; d2[1] = 2; break;
    case OG_T: off[2] = Lf->n[2] / 2; d1[2] = -1; d2[2] = -2; break;
    default: d1[2] = 1; d2[2] = 2; break;
    }
    double c1 = -3.0 / 2.0, c2 = 2.0, c3 = -0.5;
    if (loc == OG_E || loc == OG_N || loc == OG_T) {
        c3 = 0.5;
        c2 = -2.0;
        c1 = 3.0 / 2.0;
    }
    for (int64_t i = 0; i < cfs[0]; i++)
        for (int64_t j = 0; j < cfs[1]; j++)
            for (int64_t k = 0; k < cfs[2]; k++) {
                const int64_t a = 2 * (i + off[0]), b = 2 * (j + off[1]), c = 2 * (k + off[2]);
                g_coarse[IX(&GC, i, j, k)] -= (c1 * u_fine[IX(Lf, a, b, c)] +
                                               c2 * u_fine[IX(Lf, a + d1[0], b + d1[1], c + d1[2])] +
                                               c3 * u_fine[IX(Lf, a + d2[0], b + d2[1], c + d2[2])]) / h_fine;
            }
}

/* Reference: include/devicearray.h:198-240 — max |a| over the shape */
double og_norm_inf(const og_layout* L, const double* a) {
    double res = 0.0;
#pragma omp parallel for schedule(static) reduction(max : res)
    for (int64_t i = 0; i < L->n[0]; i++)
        for (int64_t j = 0; j < L->n[1]; j++)
            for (int64_t k = 0; k < L->n[2]; k++) {
                const double v = fabs(a[IX(L, i, j, k)]);
                res = v > res ? v : res;
            }
    return res;
}

/* include/devicearray.h:161-196 — a += b over the shape */
static void add_shape(const og_layout* L, double* a, const double* b) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L->n[0]; i++)
        for (int64_t j = 0; j < L->n[1]; j++)
            for (int64_t k = 0; k < L->n[2]; k++) {
                const int64_t c = IX(L, i, j, k);
                a[c] += b[c];
            }
}

/* ------------------------------------------------------------------ solver */

typedef struct og_level {
    og_layout L;
    double h;
    double *u, *p, *f, *r; /* p == u for Gauss-Seidel (domain.h:114-119) */
    double* g[6];
    og_layout GL[6];
    int32_t type[6]; /* effective boundary object type */
} og_level;

struct og_solver {
    og_params P;
    int nl;
    og_level* lv;
    int alternate; /* GaussSeidel red/black toggle, gaussseidel.h:18,93 */
};

static double* zalloc(int64_t n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

/* include/libpoisson.h:113-156, include/domain.h:100-193,
 * src/domainsettings.cpp:12-29 */
og_solver* og_create(const og_params* p) {
    if (p->levels < 1) return NULL;
    og_solver* s = (og_solver*)calloc(1, sizeof(og_solver));
    s->P = *p;
    s->nl = (int)p->levels;
    s->lv = (og_level*)calloc((size_t)s->nl, sizeof(og_level));
    s->alternate = 1;
    if (p->threads > 0) omp_set_num_threads(p->threads);
    int64_t halo[6];
    for (int q = 0; q < 6; q++) halo[q] = p->bc[q] != OG_DIRICHLET;
    for (int l = 0; l < s->nl; l++) {
        og_level* v = &s->lv[l];
        int64_t n[3];
        for (int a = 0; a < 3; a++) {
            n[a] = l == 0 ? p->dims[a] : (p->dims[a] >> l) + 1;
            if (l > 0 && (((n[a] - 1) << l) + 1) != p->dims[a]) {
                og_destroy(s);
                return NULL; /* invalid grid: std::invalid_argument in the reference */
            }
        }
        og_layout_make(&v->L, n[0], n[1], n[2], halo);
        v->h = p->h * (double)(1L << l);
        v->u = zalloc(v->L.size);
        v->p = p->relaxation == OG_JACOBI ? zalloc(v->L.size) : v->u;
        v->f = zalloc(v->L.size);
        v->r = zalloc(v->L.size);
        for (int q = 0; q < 6; q++) {
            int64_t fs[3];
            face_shape(q, n, fs);
            og_layout_make(&v->GL[q], fs[0], fs[1], fs[2], NULL);
            v->g[q] = zalloc(v->GL[q].size);
            v->type[q] = p->bc[q];
        }
        if (p->mode == OG_MODE_REFERENCE) {
            /* domain.h:139-145 (bottom always Dirichlet), :191-192 (north and
             * south always Dirichlet) */
            v->type[OG_N] = OG_DIRICHLET;
            v->type[OG_S] = OG_DIRICHLET;
            v->type[OG_B] = OG_DIRICHLET;
        }
    }
    if (p->problem >= 0) {
        /* include/libpoisson.h:159-171: level 0 from the analytic problem */
        og_level* v = &s->lv[0];
        const double x0 = p->origin[0], y0 = p->origin[1], z0 = p->origin[2], h = v->h;
        for (int q = 0; q < 6; q++) {
            /* boundary.h:82-111 (init_by_fun) with dirichlet.h:61 / neumann.h:46-61 */
            double off[3] = {0, 0, 0};
            if (q == OG_T) off[2] = (double)(v->L.n[2] - 1);
            if (q == OG_N) off[1] = (double)(v->L.n[1] - 1);
            if (q == OG_E) off[0] = (double)(v->L.n[0] - 1);
            const og_layout* G = &v->GL[q];
            for (int64_t i = 0; i < G->n[0]; i++)
                for (int64_t j = 0; j < G->n[1]; j++)
                    for (int64_t k = 0; k < G->n[2]; k++) {
                        const double x = x0 + (((double)i) + off[0]) * h;
                        const double y = y0 + (((double)j) + off[1]) * h;
                        const double z = z0 + (((double)k) + off[2]) * h;
                        double val;
                        if (v->type[q] == OG_DIRICHLET)
                            val = og_ufun(p->problem, x, y, z);
                        else if (q == OG_E || q == OG_W)
                            val = og_dudx(p->problem, x, y, z);
                        else if (q == OG_N || q == OG_S)
                            val = og_dudy(p->problem, x, y, z);
                        else
                            val = og_dudz(p->problem, x, y, z);
                        v->g[q][IX(G, i, j, k)] = val;
                    }
        }
        /* domain.h:329-342 (init_f) */
        for (int64_t i = 0; i < v->L.n[0]; i++)
            for (int64_t j = 0; j < v->L.n[1]; j++)
                for (int64_t k = 0; k < v->L.n[2]; k++)
                    v->f[IX(&v->L, i, j, k)] =
                        og_ffun(p->problem, x0 + ((double)i) * h, y0 + ((double)j) * h, z0 + ((double)k) * h);
    }
    return s;
}

void og_destroy(og_solver* s) {
    if (!s) return;
    for (int l = 0; l < s->nl; l++) {
        og_level* v = &s->lv[l];
        free(v->u);
        if (v->p != v->u) free(v->p);
        free(v->f);
        free(v->r);
        for (int q = 0; q < 6; q++) free(v->g[q]);
    }
    free(s->lv);
    free(s);
}

/* include/domain.h:319-326 (write_bc_to), order N,S,E,W,T,B */
static void write_bc_to(og_level* v, double* u) {
    for (int q = 0; q < 6; q++) {
        if (v->type[q] == OG_DIRICHLET)
            og_bc_dirichlet(&v->L, u, q, v->g[q]);
        else
            og_bc_neumann(&v->L, u, q, v->g[q], v->h);
    }
}

/* include/domain.h:246-263 */
void og_to_device(og_solver* s) {
    for (int l = 0; l < s->nl; l++) {
        og_level* v = &s->lv[l];
        write_bc_to(v, v->u);
        if (v->p != v->u) write_bc_to(v, v->p);
    }
}

/* Boundary::update for all faces, order E,W,N,S,T,B (relaxation.h:35-52,
 * residual.h:26-34); only Neumann faces act (dirichlet.h:95-99). */
static void update_bcs(og_level* v, double* target) {
    static const int order[6] = {OG_E, OG_W, OG_N, OG_S, OG_T, OG_B};
    for (int t = 0; t < 6; t++) {
        const int q = order[t];
        if (v->type[q] == OG_NEUMANN) og_bc_neumann(&v->L, target, q, v->g[q], v->h);
    }
}

/* relaxation.h:78-83 with Dirichlet::is_non_eliminated() == true */
static void level_box(const og_level* v, int64_t box[6]) {
    box[0] = v->type[OG_W] == OG_DIRICHLET;
    box[1] = v->L.n[0] - (v->type[OG_E] == OG_DIRICHLET);
    box[2] = v->type[OG_S] == OG_DIRICHLET;
    box[3] = v->L.n[1] - (v->type[OG_N] == OG_DIRICHLET);
    box[4] = v->type[OG_B] == OG_DIRICHLET;
    box[5] = v->L.n[2] - (v->type[OG_T] == OG_DIRICHLET);
}

/* include/relaxation.h:19-94 */
static void relax(og_solver* s, int l, double omega, int64_t nsmooth) {
    og_level* v = &s->lv[l];
    const int gs = s->P.relaxation == OG_GAUSS_SEIDEL;
    const int64_t mult = gs ? 2 : 1;
    int64_t box[6];
    level_box(v, box);
    for (int64_t it = 0; it < mult * nsmooth; it++) {
        if (!gs) {
            double* t = v->u; /* domain.h:349-354 swap_u */
            v->u = v->p;
            v->p = t;
        }
        update_bcs(v, gs ? v->u : v->p);
        if (gs) {
            og_gs_half(s->P.mode, &v->L, v->u, v->f, box, omega, v->h, s->alternate ? 0 : 1);
            s->alternate = !s->alternate;
        } else {
            og_jacobi(s->P.mode, &v->L, v->u, v->p, v->f, box, omega, v->h);
        }
    }
}

/* include/residual.h:18-81 */
static void residual(og_solver* s, int l) {
    og_level* v = &s->lv[l];
    int64_t box[6];
    update_bcs(v, v->u);
    level_box(v, box);
    og_residual(s->P.mode, &v->L, v->u, v->f, v->r, box, v->h);
}

static void restrict_to(og_solver* s, const og_layout* Li, const double* in, const og_layout* Lo, double* out) {
    if (s->P.restriction == OG_FULL_WEIGHTING)
        og_restrict_fw(Li, in, Lo, out);
    else
        og_restrict_inject(Li, in, Lo, out);
}

/* vcycle.h:34-76 / fcycle.h:130-172: defect, restriction, coarse zero,
 * boundary data of the coarse level (order N,S,E,W,T,B). */
static void go_down(og_solver* s, int l) {
    og_level* v = &s->lv[l];
    og_level* c = &s->lv[l + 1];
    residual(s, l);
    restrict_to(s, &v->L, v->r, &c->L, c->f);
    memset(c->u, 0, sizeof(double) * (size_t)c->L.size);
    for (int q = 0; q < 6; q++) {
        if (c->type[q] == OG_DIRICHLET || s->P.mode == OG_MODE_CONSISTENT) {
            /* dirichlet.h:101-107; consistent mode: homogeneous coarse data */
            memset(c->g[q], 0, sizeof(double) * (size_t)c->GL[q].size);
        } else {
            og_bc_restrict_neumann(s->P.restriction, &v->L, v->u, q, v->g[q], c->g[q], v->h);
        }
    }
}

/* vcycle.h:81-94: prolong the coarse correction into r_l (scratch), u += r */
static void go_up(og_solver* s, int l) {
    og_level* v = &s->lv[l];
    og_level* c = &s->lv[l + 1];
    og_prolong_trilinear(&c->L, c->u, &v->L, v->r);
    add_shape(&v->L, v->u, v->r);
}

/* include/vcycle.h:14-98 (coarsest level: nsmooth passes, or coarse_sweeps in
 * consistent mode) */
static void vcycle(og_solver* s, int l, double omega, int64_t nsmooth) {
    if (l >= s->nl - 1) {
        const int64_t n = (s->P.mode == OG_MODE_CONSISTENT && s->P.coarse_sweeps > 0) ? s->P.coarse_sweeps : nsmooth;
        relax(s, l, omega, n);
        return;
    }
    relax(s, l, omega, nsmooth);
    go_down(s, l);
    vcycle(s, l + 1, omega, nsmooth);
    go_up(s, l);
    relax(s, l, omega, nsmooth);
}

/* include/fcycle.h:13-97 */
static void fcycle(og_solver* s, int l, double omega, int64_t nsmooth) {
    relax(s, l, omega, nsmooth);
    go_down(s, l);
    if (l < s->nl - 2) fcycle(s, l + 1, omega, nsmooth);
    vcycle(s, l + 1, omega, nsmooth);
    go_up(s, l);
    if (l == 0) vcycle(s, 0, omega, nsmooth);
}

void og_cycle(og_solver* s, int cycle, int64_t nsmooth, double omega) {
    if (omega < 0) omega = s->P.relaxation == OG_JACOBI ? 6.0 / 7.0 : 8.0 / 9.0;
    if (cycle == OG_VCYCLE)
        vcycle(s, 0, omega, nsmooth);
    else if (cycle == OG_FCYCLE)
        fcycle(s, 0, omega, nsmooth);
    /* "relaxation": no work per iteration (libpoisson.h:304-308) */
}

void og_relax_level(og_solver* s, int level, double omega, int64_t nsmooth) { relax(s, level, omega, nsmooth); }
void og_residual_level(og_solver* s, int level) { residual(s, level); }

double og_relres(og_solver* s) {
    og_level* v = &s->lv[0];
    residual(s, 0);
    return og_norm_inf(&v->L, v->r) / og_norm_inf(&v->L, v->f);
}

/* include/libpoisson.h:237-357 (wall-time limit not modelled). */
int64_t og_solve(og_solver* s, int cycle, int64_t nsmooth, double omega, int64_t maxiter, int64_t miniter,
                 double tol, double* relres_out, double* final_relres) {
    og_level* v = &s->lv[0];
    const double fnorm = og_norm_inf(&v->L, v->f);
    residual(s, 0);
    double rel_res = og_norm_inf(&v->L, v->r) / fnorm;
    int64_t iter;
    for (iter = 0; iter < maxiter; iter++) {
        og_cycle(s, cycle, nsmooth, omega);
        residual(s, 0);
        double norm = og_norm_inf(&v->L, v->r);
        norm /= fnorm;
        if (relres_out) relres_out[iter] = norm;
        if (s->P.mode == OG_MODE_CONSISTENT && norm <= tol) {
            rel_res = norm;
            break;
        }
        if (iter >= miniter) {
            if (norm > 2.0 * rel_res) {
                rel_res = norm;
                break;
            } else if (s->P.mode == OG_MODE_REFERENCE && fabs(norm - rel_res) < tol) {
                rel_res = norm;
                break;
            }
        }
        rel_res = norm;
    }
    if (final_relres) *final_relres = rel_res;
    return iter;
}

double* og_array(og_solver* s, int level, int which) {
    og_level* v = &s->lv[level];
    switch (which) {
    case 'u': return v->u;
    case 'p': return v->p;
    case 'f': return v->f;
    default: return v->r;
    }
}
const og_layout* og_array_layout(og_solver* s, int level) { return &s->lv[level].L; }
double* og_face(og_solver* s, int level, int loc) { return s->lv[level].g[loc]; }
const og_layout* og_face_layout(og_solver* s, int level, int loc) { return &s->lv[level].GL[loc]; }
int32_t og_face_type(og_solver* s, int level, int loc) { return s->lv[level].type[loc]; }
double og_level_h(og_solver* s, int level) { return s->lv[level].h; }

/* ------------------------------------------ FP16 storage, FP32 arithmetic */

/* binary16 bits -> float (exact) */
float og_h2f(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 31u, m = h & 1023u;
    uint32_t b;
    if (e == 31) {
        b = sign | 0x7f800000u | (m << 13);
    } else if (e == 0) {
        float v = ldexpf((float)m, -24); /* subnormal (or zero): m 2^-24, exact */
        memcpy(&b, &v, 4);
        b |= sign;
    } else {
        b = sign | ((e + 112u) << 23) | (m << 13); /* exponent bias 15 -> 127 */
    }
    float out;
    memcpy(&out, &b, 4);
    return out;
}

/* float -> binary16, round to nearest, ties to even */
uint16_t og_f2h(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    const uint16_t sign = (uint16_t)((b >> 16) & 0x8000u);
    const uint32_t a = b & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);           /* NaN */
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);          /* >= 65520 rounds to infinity */
    if (a < 0x38800000u) {                                            /* below 2^-14: subnormal result */
        float v;
        memcpy(&v, &a, 4);
        return (uint16_t)(sign | (uint16_t)nearbyintf(v * 16777216.0f)); /* v 2^24 is exact; RNE to an integer */
    }
    const uint32_t e = (a >> 23) - 112u; /* 1..30 */
    const uint32_t m = a & 0x7fffffu;
    uint32_t hb = (e << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (hb & 1u))) hb++; /* a carry moves into the exponent */
    return (uint16_t)(sign | hb);
}

float og_h_scale(double norm, int inverse) {
    if (!(norm > 0.0) || !(norm < 1e300)) return 1.0f;
    int e;
    (void)frexp(norm, &e);
    e = MIN(120, MAX(-120, e));
    return ldexpf(1.0f, inverse ? e : -e);
}

#define HV(a, c) og_h2f((a)[c])

/* consistent-mode gs_point on binary16 storage; the constants as the FP32
 * smoother forms them from (float)omega and h */
void og_h_gs_half(const og_layout* L, uint16_t* u, const uint16_t* f, const int64_t box[6], double omega, double h,
                  int colour) {
    const float omega_sixth = (float)((double)(float)omega / 6.0);
    const float hsq = (float)(h * h);
    const int64_t s0 = L->s[0], s1 = L->s[1];
#pragma omp parallel for schedule(static)
    for (int64_t i = box[0]; i < box[1]; i++)
        for (int64_t j = box[2]; j < box[3]; j++)
            for (int64_t k = box[4]; k < box[5]; k++) {
                if ((int)((i + j + k) & 1) != colour) continue;
                const int64_t c = IX(L, i, j, k);
                const float uc = HV(u, c);
                const float s = ((HV(u, c - s0) - uc) + (HV(u, c + s0) - uc)) +
                                ((HV(u, c - s1) - uc) + (HV(u, c + s1) - uc)) + ((HV(u, c - 1) - uc) + (HV(u, c + 1) - uc));
                u[c] = og_f2h(fmaf(omega_sixth, fmaf(-hsq, HV(f, c), s), uc));
            }
}

double og_h_residual(const og_layout* L, const uint16_t* u, const uint16_t* f, uint16_t* r, const int64_t box[6],
                     double h) {
    const float inv_h2 = (float)(1.0 / (h * h));
    const int64_t s0 = L->s[0], s1 = L->s[1];
#pragma omp parallel for schedule(static)
    for (int64_t i = box[0]; i < box[1]; i++)
        for (int64_t j = box[2]; j < box[3]; j++)
            for (int64_t k = box[4]; k < box[5]; k++) {
                const int64_t c = IX(L, i, j, k);
                const float uc = HV(u, c);
                const float s = ((HV(u, c - s0) - uc) + (HV(u, c + s0) - uc)) +
                                ((HV(u, c - s1) - uc) + (HV(u, c + s1) - uc)) + ((HV(u, c - 1) - uc) + (HV(u, c + 1) - uc));
                r[c] = og_f2h(fmaf(-inv_h2, s, HV(f, c)));
            }
    double m = 0.0;
    for (int64_t i = 0; i < L->n[0]; i++)
        for (int64_t j = 0; j < L->n[1]; j++)
            for (int64_t k = 0; k < L->n[2]; k++) m = MAX(m, fabs((double)HV(r, IX(L, i, j, k))));
    return m;
}

/* og_restrict_fw's weights and term order with the running sum held in
 * float between the three planes, as the FP32 kernels keep it */
void og_h_restrict_fw(const og_layout* Li, const void* in, int in_f32, const og_layout* Lo, uint16_t* out,
                      double norm) {
    const int64_t n0 = Li->n[0], n1 = Li->n[1], n2 = Li->n[2];
    const float sigma = og_h_scale(norm, 0);
    const float* in32 = (const float*)in;
    const uint16_t* in16 = (const uint16_t*)in;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Lo->n[0]; i++)
        for (int64_t j = 0; j < Lo->n[1]; j++)
            for (int64_t k = 0; k < Lo->n[2]; k++) {
                const int64_t ii = MIN(n0 - 1, 2 * i), jj = MIN(n1 - 1, 2 * j), kk = MIN(n2 - 1, 2 * k);
                const int64_t im = MAX(0, ii - 1), ip = MIN(n0 - 1, ii + 1);
                const int64_t jm = MAX(0, jj - 1), jp = MIN(n1 - 1, jj + 1);
                const int64_t km = MAX(0, kk - 1), kp = MIN(n2 - 1, kk + 1);
#define A(a, b, c) (in_f32 ? in32[IX(Li, a, b, c)] : og_h2f(in16[IX(Li, a, b, c)]))
                float tmp = (float)(1.0 * A(im, jm, km) + 2.0 * A(im, jj, km) + 1.0 * A(im, jp, km) +
                                    2.0 * A(ii, jm, km) + 4.0 * A(ii, jj, km) + 2.0 * A(ii, jp, km) +
                                    1.0 * A(ip, jm, km) + 2.0 * A(ip, jj, km) + 1.0 * A(ip, jp, km));
                tmp = (float)((double)tmp + (2.0 * A(im, jm, kk) + 4.0 * A(im, jj, kk) + 2.0 * A(im, jp, kk) +
                                             4.0 * A(ii, jm, kk) + 8.0 * A(ii, jj, kk) + 4.0 * A(ii, jp, kk) +
                                             2.0 * A(ip, jm, kk) + 4.0 * A(ip, jj, kk) + 2.0 * A(ip, jp, kk)));
                tmp = (float)((double)tmp + (1.0 * A(im, jm, kp) + 2.0 * A(im, jj, kp) + 1.0 * A(im, jp, kp) +
                                             2.0 * A(ii, jm, kp) + 4.0 * A(ii, jj, kp) + 2.0 * A(ii, jp, kp) +
                                             1.0 * A(ip, jm, kp) + 2.0 * A(ip, jj, kp) + 1.0 * A(ip, jp, kp)));
#undef A
                const float v = (float)((double)tmp / 64.0);
                out[IX(Lo, i, j, k)] = og_f2h(v * sigma);
            }
}

/* 0.5*(a+b): the float sum, halved */
static inline float half_sum_f(float a, float b) { return (float)(0.5 * (double)(a + b)); }

/* trilinear value at fine (i,j,k) from the coarse values around it,
 * composed as og_prolong_trilinear's x, y and z passes compose it */
static float tri_xy_f(const og_layout* Lc, const uint16_t* c, int64_t a, int64_t b, int64_t cc, int pi, int pj) {
#define C(x, y, z) og_h2f(c[IX(Lc, x, y, z)])
    if (!pi && !pj) return C(a, b, cc);
    if (pi && !pj) return half_sum_f(C(a, b, cc), C(a + 1, b, cc));
    if (!pi && pj) return half_sum_f(C(a, b, cc), C(a, b + 1, cc));
    return half_sum_f(half_sum_f(C(a, b, cc), C(a + 1, b, cc)), half_sum_f(C(a, b + 1, cc), C(a + 1, b + 1, cc)));
#undef C
}

void og_h_prolong_add(const og_layout* Lc, const uint16_t* coarse, const og_layout* Lf, void* fine, int fine_f32,
                      double norm) {
    const float inv = og_h_scale(norm, 1);
    float* f32 = (float*)fine;
    uint16_t* f16 = (uint16_t*)fine;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Lf->n[0]; i++)
        for (int64_t j = 0; j < Lf->n[1]; j++)
            for (int64_t k = 0; k < Lf->n[2]; k++) {
                const int pi = (int)(i & 1), pj = (int)(j & 1), pk = (int)(k & 1);
                const float lo = tri_xy_f(Lc, coarse, i >> 1, j >> 1, k >> 1, pi, pj);
                const float v = (pk ? half_sum_f(lo, tri_xy_f(Lc, coarse, i >> 1, j >> 1, (k >> 1) + 1, pi, pj)) : lo) *
                                inv;
                const int64_t c = IX(Lf, i, j, k);
                if (fine_f32) f32[c] = f32[c] + v;
                else f16[c] = og_f2h(og_h2f(f16[c]) + v);
            }
}
#undef HV

void og_f2h_n(const float* x, uint16_t* out, int64_t n) {
    for (int64_t t = 0; t < n; t++) out[t] = og_f2h(x[t]);
}
void og_h2f_n(const uint16_t* x, float* out, int64_t n) {
    for (int64_t t = 0; t < n; t++) out[t] = og_h2f(x[t]);
}

/* ------------------------------------------------ hi + lo refinement iterate
 * (this project's addition; no counterpart in the reference) */

/* Knuth's TwoSum (The Art of Computer Programming vol. 2, 4.2.2):
 * s = fl(a + b), a + b = s + *e exactly (built with -ffp-contract=off) */
static double og_two_sum(double a, double b, double* e) {
    const double s = a + b;
    const double bb = s - a;
    *e = (a - (s - bb)) + (b - bb);
    return s;
}

void og_hilo_update(const og_layout* L, double* hi, float* lo, const float* e) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L->n[0]; i++)
        for (int64_t j = 0; j < L->n[1]; j++)
            for (int64_t k = 0; k < L->n[2]; k++) {
                const int64_t c = IX(L, i, j, k);
                double err;
                hi[c] = og_two_sum(hi[c], (double)e[c] + (double)lo[c], &err);
                lo[c] = (float)err;
            }
}

/* (hi, lo) of the neighbour of in-shape point (i,j,k) one step along `axis`
 * in direction `dir`: stored, or the Neumann ghost formed from the mirror */
static void og_hilo_nb(const og_layout* L, const double* hi, const float* lo, const double* const g[6], double two_h,
                       int64_t i, int64_t j, int64_t k, int axis, int dir, double* vh, float* vl) {
    int64_t a[3] = {i, j, k};
    a[axis] += dir;
    if (a[axis] >= 0 && a[axis] < L->n[axis]) {
        const int64_t c = IX(L, a[0], a[1], a[2]);
        *vh = hi[c];
        *vl = lo[c];
        return;
    }
    int face;
    int64_t gi;
    if (axis == 0) {
        face = dir > 0 ? OG_E : OG_W;
        gi = j * L->n[2] + k; /* (1, ny, nz) */
    } else if (axis == 1) {
        face = dir > 0 ? OG_N : OG_S;
        gi = i * L->n[2] + k; /* (nx, 1, nz) */
    } else {
        face = dir > 0 ? OG_T : OG_B;
        gi = i * L->n[1] + j; /* (nx, ny, 1) */
    }
    a[axis] -= 2 * dir; /* mirror */
    const int64_t m = IX(L, a[0], a[1], a[2]);
    const double p = (dir > 0 ? 1.0 : -1.0) * two_h * g[face][gi];
    double err;
    *vh = og_two_sum(p, hi[m], &err);
    *vl = (float)((double)lo[m] + err);
}

double og_hilo_residual(const og_layout* L, const double* hi, const float* lo, const double* f, double* r,
                        const double* const g[6], const int64_t box[6], double h) {
    const double two_h = 2.0 * h, inv_h2 = 1.0 / (h * h);
    double vmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : vmax)
    for (int64_t i = box[0]; i < box[1]; i++)
        for (int64_t j = box[2]; j < box[3]; j++)
            for (int64_t k = box[4]; k < box[5]; k++) {
                const int64_t c = IX(L, i, j, k);
                double nh[6];
                float nl[6];
                /* W, E, S, N, B, T */
                for (int q = 0; q < 6; q++)
                    og_hilo_nb(L, hi, lo, g, two_h, i, j, k, q / 2, (q & 1) ? 1 : -1, &nh[q], &nl[q]);
                const double hc = hi[c], lc = (double)lo[c];
                const double sh = ((nh[0] - hc) + (nh[1] - hc)) + ((nh[2] - hc) + (nh[3] - hc)) +
                                  ((nh[4] - hc) + (nh[5] - hc));
                const double sl = (((double)nl[0] - lc) + ((double)nl[1] - lc)) +
                                  (((double)nl[2] - lc) + ((double)nl[3] - lc)) +
                                  (((double)nl[4] - lc) + ((double)nl[5] - lc));
                r[c] = fma(-inv_h2, sh + sl, f[c]);
                vmax = MAX(vmax, fabs(r[c]));
            }
    for (int64_t i = 0; i < L->n[0]; i++)
        for (int64_t j = 0; j < L->n[1]; j++)
            for (int64_t k = 0; k < L->n[2]; k++) {
                if (i >= box[0] && i < box[1] && j >= box[2] && j < box[3] && k >= box[4] && k < box[5]) continue;
                vmax = MAX(vmax, fabs(r[IX(L, i, j, k)]));
            }
    return vmax;
}
