/*
 * CPU ORACLE — test infrastructure only.
 *
 * A plain-C, FP64 restatement of the reference's geometric-multigrid solve
 * (eljens/multigrid-poisson, include/ headers) used as the checker for the HIP
 * library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it; the product never links, calls or falls back to it.
 *
 * Two modes, selected per solver:
 *   OG_MODE_REFERENCE  — the reference's behaviour, quirks included, with the
 *                        reference's exact arithmetic sequence (bit-identical
 *                        to the reference built without FP contraction);
 *   OG_MODE_CONSISTENT — the six requested faces honoured, homogeneous
 *                        coarse-grid Neumann data, (i+j+k) colouring,
 *                        cancellation-aware stencils with explicit fma(), and a
 *                        ||r||/||f|| <= tol stop.
 *
 * How each function is pinned (details in tests/test_oracle.py):
 *   - reference mode: against golden vectors recorded from the reference
 *     itself (tests/golden/), bit for bit;
 *   - consistent mode: closed forms (exactness of the 7-point stencil on
 *     cubics, Jacobi eigen-damping, ghost-node Neumann identity), a dense
 *     direct solve of the discrete system on tiny grids, and second-order
 *     convergence against the analytic solution.
 *
 * Layout (reference include/array.h:84-94, 130-141): halo-padded C order,
 * x slowest, z contiguous; idx(i,j,k) = (i+hW)*s0 + (j+hS)*s1 + (k+hB).
 * Face and halo arrays are indexed in Location_t order (include/boundary.h:12):
 * NORTH, SOUTH, EAST, WEST, TOP, BOTTOM.
 */
#ifndef MG_ORACLE_H
#define MG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OG_N = 0, OG_S = 1, OG_E = 2, OG_W = 3, OG_T = 4, OG_B = 5 };
enum { OG_DIRICHLET = 0, OG_NEUMANN = 1 };
enum { OG_MODE_REFERENCE = 0, OG_MODE_CONSISTENT = 1 };
enum { OG_JACOBI = 0, OG_GAUSS_SEIDEL = 1 };
enum { OG_INJECTION = 0, OG_FULL_WEIGHTING = 1 };
enum { OG_RELAXATION = 0, OG_VCYCLE = 1, OG_FCYCLE = 2 };

typedef struct og_layout {
    int64_t n[3];     /* points per axis, halos excluded */
    int64_t halo[6];  /* Location_t order */
    int64_t s[3];     /* element strides, s[2] == 1 */
    int64_t size;     /* elements incl. halos */
} og_layout;

typedef struct og_params {
    int64_t dims[3];
    int64_t levels;
    double h;           /* finest spacing (isotropic, src/parser.cpp:48) */
    double origin[3];
    int32_t bc[6];      /* requested BC per face, Location_t order */
    int32_t mode;       /* OG_MODE_* */
    int32_t relaxation; /* OG_JACOBI / OG_GAUSS_SEIDEL */
    int32_t restriction;/* OG_INJECTION / OG_FULL_WEIGHTING */
    int32_t problem;    /* analytic problem id for init(); <0: init_zero() */
    int32_t threads;    /* OpenMP threads, 0 = runtime default */
    int32_t coarse_sweeps; /* consistent mode: smoothing passes on the coarsest
                              level per visit (0 = nsmooth, as the reference) */
} og_params;

/* ---- layout helpers ---- */
void og_layout_make(og_layout* L, int64_t nx, int64_t ny, int64_t nz, const int64_t halo[6]);

/* ---- analytic problems (same definitions the drivers use) ---- */
double og_ufun(int problem, double x, double y, double z);
double og_ffun(int problem, double x, double y, double z);
double og_dudx(int problem, double x, double y, double z);
double og_dudy(int problem, double x, double y, double z);
double og_dudz(int problem, double x, double y, double z);

/* ---- single operators on caller buffers (layouts given) ---- */
/* box = {xmin,xmax,ymin,ymax,zmin,zmax} */
void og_residual(int mode, const og_layout* L, const double* u, const double* f, double* r,
                 const int64_t box[6], double h);
void og_gs_half(int mode, const og_layout* L, double* u, const double* f, const int64_t box[6],
                double omega, double h, int colour);
void og_jacobi(int mode, const og_layout* L, double* u, const double* v, const double* f,
               const int64_t box[6], double omega, double h);
void og_restrict_inject(const og_layout* Li, const double* in, const og_layout* Lo, double* out);
void og_restrict_fw(const og_layout* Li, const double* in, const og_layout* Lo, double* out);
void og_prolong_trilinear(const og_layout* Li, const double* in, const og_layout* Lo, double* out);
void og_bc_dirichlet(const og_layout* L, double* u, int loc, const double* g);
void og_bc_neumann(const og_layout* L, double* u, int loc, const double* g, double h);
/* coarse g = R(g_fine) - one-sided derivative of u_fine / h_fine (reference mode) */
void og_bc_restrict_neumann(int restriction, const og_layout* Lf, const double* u_fine, int loc,
                            const double* g_fine, double* g_coarse, double h_fine);
double og_norm_inf(const og_layout* L, const double* a);

/* ---- whole solver (one device, PoissonSolver<double,R,Trilinear,S>) ---- */
typedef struct og_solver og_solver;
og_solver* og_create(const og_params* p);
void og_destroy(og_solver* s);
void og_to_device(og_solver* s);          /* write_bc_to(u) (+uprev) on level 0..L-1 */
/* one cycle (reference Vcycle/Fcycle from level 0) */
void og_cycle(og_solver* s, int cycle, int64_t nsmooth, double omega);
/* reference solve loop; relres_out receives the relative residual after each
 * executed cycle (at most maxiter values); returns solve_iterations(). */
int64_t og_solve(og_solver* s, int cycle, int64_t nsmooth, double omega, int64_t maxiter,
                 int64_t miniter, double tol, double* relres_out, double* final_relres);
/* one level's Relaxation::relax (boundary refresh + sweeps) and residual() */
void og_relax_level(og_solver* s, int level, double omega, int64_t nsmooth);
void og_residual_level(og_solver* s, int level);
/* residual of level 0 (with boundary refresh) and its inf-norm / ||f||inf */
double og_relres(og_solver* s);
/* access: which = 'u','f','r','p' (uprev), face arrays 'g'+loc via og_face */
double* og_array(og_solver* s, int level, int which);
const og_layout* og_array_layout(og_solver* s, int level);
double* og_face(og_solver* s, int level, int loc);
const og_layout* og_face_layout(og_solver* s, int level, int loc);
int32_t og_face_type(og_solver* s, int level, int loc);
double og_level_h(og_solver* s, int level);

/* ---- FP16 storage, FP32 arithmetic (the coarse levels of an FP32 cycle;
 * this project's addition, the reference has one type T throughout).
 * Arrays hold IEEE binary16 bit patterns; values are loaded as float,
 * combined in float with the consistent-mode forms (fmaf), and rounded to
 * binary16 (nearest, ties to even) when stored.  Consistent mode only:
 * colour = parity of i+j+k, boxes strictly inside the shape.
 * Pinned by: og_f2h/og_h2f against numpy's binary16 conversion (every
 * binary16 value, every midpoint between neighbours, random floats); each
 * operator against the FP64 operator above on inputs where every float
 * operation is exact (so the two must agree after rounding), and against it
 * within the binary16 rounding bound on random inputs. */
uint16_t og_f2h(float x);
float og_h2f(uint16_t h);
void og_f2h_n(const float* x, uint16_t* out, int64_t n);
void og_h2f_n(const uint16_t* x, float* out, int64_t n);
/* 2^-e (inverse: 2^e) for norm = m 2^e, m in [0.5,1); 1 for norm <= 0 */
float og_h_scale(double norm, int inverse);
void og_h_gs_half(const og_layout* L, uint16_t* u, const uint16_t* f, const int64_t box[6], double omega, double h,
                  int colour);
/* r = f - L u on the box; returns max |r| of the stored r over its shape */
double og_h_residual(const og_layout* L, const uint16_t* u, const uint16_t* f, uint16_t* r, const int64_t box[6],
                     double h);
/* out = sigma(norm) * FW(in), in binary16 (in_f32 = 0) or float (in_f32 = 1) */
void og_h_restrict_fw(const og_layout* Li, const void* in, int in_f32, const og_layout* Lo, uint16_t* out,
                      double norm);
/* fine += P(coarse) / sigma(norm) over the fine shape, fine binary16 or float */
void og_h_prolong_add(const og_layout* Lc, const uint16_t* coarse, const og_layout* Lf, void* fine, int fine_f32,
                      double norm);

/* ---- the iterate of mixed-precision refinement held as hi + lo (FP64 hi,
 * FP32 lo; this project's addition, no counterpart in the reference).
 * Consistent mode only.  Update: hi' = fl(hi + fl(e + lo)) by Knuth's
 * error-free TwoSum, lo' = (float) of its rounding error.  Residual of hi + lo:
 * r = fma(-1/h^2, S(hi) + S(lo), f) with S the consistent-mode sum of
 * neighbour differences (x pair, y pair, z pair), the lo differences in FP64.
 * A neighbour beyond a Neumann face is the ghost p + mirror, p =
 * (sign*2h)*g as include/neumann.h:112-130 forms it, hi = fl(p + hi(mirror))
 * and lo = (float)(lo(mirror) + rounding error of that sum).
 * Pinned by: exact rational arithmetic (Python fractions) on inputs where
 * every FP64 operation is exact, and by the TwoSum identity hi' + err =
 * hi + fl(e + lo) on random inputs.
 * g[q]: face data of Neumann faces (layouts as in mgp.h), NULL elsewhere. */
void og_hilo_update(const og_layout* L, double* hi, float* lo, const float* e);
/* r on the box (r elsewhere untouched); returns max |r| over the box and r's
 * stored values outside it */
double og_hilo_residual(const og_layout* L, const double* hi, const float* lo, const double* f, double* r,
                        const double* const g[6], const int64_t box[6], double h);

#ifdef __cplusplus
}
#endif
#endif
