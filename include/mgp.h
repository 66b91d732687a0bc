/*
 * mgp — C ABI of the MI355X (gfx950) geometric-multigrid library.
 *
 * Every entry point takes plain pointers and sizes, returns 0 on success and
 * otherwise a hipError_t value or one of the MGP_ERR_* codes below, throws
 * nothing, and keeps no pointer after it returns.  Work is enqueued on the
 * `stream` argument (a hipStream_t; NULL = the library's default stream) and
 * is asynchronous unless stated.
 *
 * Layout of every 3-D array (the reference's Array<T>, include/array.h:84-141
 * of eljens/multigrid-poisson): halo-padded, C order, x slowest and z
 * contiguous; element (i,j,k) of the shape sits at
 *     (i + halo[WEST])*stride[0] + (j + halo[SOUTH])*stride[1] + (k + halo[BOTTOM])
 * with i in [0,shape[0]) etc.  Face arrays of boundary data are dense arrays
 * without halos of shape (nx,ny,1) for TOP/BOTTOM, (1,ny,nz) for EAST/WEST and
 * (nx,1,nz) for NORTH/SOUTH (include/domain.h:125-190).
 *
 * Arithmetic modes:
 *   MGP_MODE_REFERENCE   the reference's expression order and types, red/black
 *                        colour = parity of the halo-inclusive linear index;
 *   MGP_MODE_CONSISTENT  cancellation-aware stencils (sums of u_nb - u_c,
 *                        explicit fma), colour = parity of i+j+k.
 * Both are exact functions of their inputs: no atomics in the arithmetic, so
 * results are reproducible bit for bit.
 */
#ifndef MGP_H
#define MGP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ enums */
#define MGP_OK 0
#define MGP_ERR_ARG (-1)    /* invalid argument: shape mismatch, bad enum, ... */
#define MGP_ERR_NODEV (-2)  /* no GPU visible */
#define MGP_ERR_STATE (-3)  /* call out of order (e.g. graph end without begin) */

enum { MGP_NORTH = 0, MGP_SOUTH = 1, MGP_EAST = 2, MGP_WEST = 3, MGP_TOP = 4, MGP_BOTTOM = 5 };
enum { MGP_DIRICHLET = 0, MGP_NEUMANN = 1 };
enum { MGP_MODE_REFERENCE = 0, MGP_MODE_CONSISTENT = 1 };
enum { MGP_INJECTION = 0, MGP_FULL_WEIGHTING = 1 };
enum { MGP_F64 = 0, MGP_F32 = 1, MGP_MIXED = 2 /* FP32 cycles in FP64 refinement (solver level) */ };

typedef struct mgp_array {
    void* data;        /* device pointer to the first element of the padded buffer */
    int64_t shape[3];  /* points per axis (x, y, z), halos excluded */
    int64_t halo[6];   /* halo width per face, order NORTH,SOUTH,EAST,WEST,TOP,BOTTOM */
    int64_t stride[3]; /* element strides of x, y, z; stride[2] == 1 */
} mgp_array;

/* Boundary faces of one level, for kernels that form Neumann ghost values
 * themselves (ghost = sign*2h*g + u(mirror), include/neumann.h:112-130). */
typedef struct mgp_faces {
    int32_t type[6];  /* effective type per face: MGP_DIRICHLET (boundary nodes fixed,
                         outside the relaxation box) or MGP_NEUMANN */
    const void* g[6]; /* device face array per Neumann face, NULL = homogeneous (g = 0) */
    double h;         /* grid spacing of the level */
} mgp_faces;

/* box = {xmin, xmax, ymin, ymax, zmin, zmax}: half-open index ranges of the
 * points a smoother or residual updates (include/relaxation.h:78-83). */

/* ----------------------------------------------------- device and memory */
int mgp_device_count(int* count);
int mgp_set_device(int device);
int mgp_malloc(void** ptr, size_t bytes);
int mgp_free(void* ptr);
int mgp_host_alloc(void** ptr, size_t bytes); /* pinned host memory */
int mgp_host_free(void* ptr);
int mgp_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int mgp_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int mgp_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int mgp_stream_create(void** stream);
int mgp_stream_destroy(void* stream);
int mgp_stream_sync(void* stream);
int mgp_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ------------------------------------------- elementwise (replaces the
 * reference's DeviceArray::init_zero/add/infinity_norm, devicearray.h:132-240) */
/* every element of the padded buffer (halos included) = value; a kernel, not a
 * memset, so it can be captured in a graph */
int mgp_fill_f64(const mgp_array* a, double value, void* stream);
int mgp_fill_f32(const mgp_array* a, double value, void* stream);
/* a += b over the shape (devicearray.h:161-196) */
int mgp_add_f64(const mgp_array* a, const mgp_array* b, void* stream);
int mgp_add_f32(const mgp_array* a, const mgp_array* b, void* stream);
/* *norm_dev = max(*norm_dev, max |a| over the shape) (devicearray.h:198-240);
 * norm_dev is a device double the caller zeroes with mgp_zero_scalars */
int mgp_norm_inf_f64(const mgp_array* a, double* norm_dev, void* stream);
int mgp_norm_inf_f32(const mgp_array* a, double* norm_dev, void* stream);
int mgp_zero_scalars(double* dev, int64_t n, void* stream);

/* ------------------------------------------- mixed precision (no counterpart
 * in the reference: its T is one type throughout) */
/* out = (float)in, out = (double)in, over the shape */
int mgp_cast_f64_f32(const mgp_array* in, const mgp_array* out, void* stream);
int mgp_cast_f32_f64(const mgp_array* in, const mgp_array* out, void* stream);
/* u += (double)e over the shape: the FP64 update of iterative refinement */
int mgp_add_f32_to_f64(const mgp_array* u, const mgp_array* e, void* stream);
/* The FP64 update and residual of one refinement iteration in one pass:
 * u_new = u + (double)e over the shape (u_new a second buffer of u's layout;
 * its halo is not written), then out = (float)(f - L u_new) on the box, the
 * Neumann ghosts formed from u_new as the boundary refresh would store them
 * (include/residual.h:18-81 after include/neumann.h:112-130), and with norm
 * max |r| over the box and r's stored values outside it.  Equals
 * mgp_add_f32_to_f64, mgp_bc_neumann_f64 on each Neumann face and
 * mgp_residual_cast_f64_f32, bit for bit, in 32 instead of 40 bytes per point.
 * No counterpart in the reference (FP64 only there). */
int mgp_update_residual_f64_f32(int mode, const mgp_array* u, const mgp_array* u_new, const mgp_array* e,
                                const mgp_array* f, const mgp_array* r, const mgp_array* out, const mgp_faces* faces,
                                const int64_t box[6], double* norm, void* stream);

/* The iterate of refinement held as hi + lo: FP64 hi plus an FP32 lo word
 * (same shape, halos and element strides as hi), about 77 significant bits,
 * so that the FP64 residual is not floored by the rounding of u to FP64
 * (12/h^2 ulp(u), about 1e-10 of ||f|| at h = 1/1024).  Consistent mode.
 * Update and residual in one pass: (u_new, lo_new) = u + lo + (double)e
 * renormalised by an error-free sum (hi' = fl(u + fl(e + lo)), lo' = the
 * rounding error as FP32), then out = (float)(f - L(u_new + lo_new)) on the
 * box, S(hi) + S(lo) summed in FP64 (S = the consistent-mode sum of
 * neighbour differences), Neumann ghosts formed from `faces` with the
 * rounding of sign*2h*g + hi(mirror) carried into lo; with norm, max |r|
 * over the box and r's stored values outside it.  u_new and lo_new are
 * second buffers (halos not written).  40 bytes per point.  No counterpart
 * in the reference (FP64 only there; its residual is include/residual.h:58-79). */
int mgp_update_residual_hilo_f64_f32(const mgp_array* u, const mgp_array* u_new, const mgp_array* lo,
                                     const mgp_array* lo_new, const mgp_array* e, const mgp_array* f,
                                     const mgp_array* r, const mgp_array* out, const mgp_faces* faces,
                                     const int64_t box[6], double* norm, void* stream);
/* r = f - L(u + lo) on the box written in FP64 (the residual of the stored
 * hi + lo iterate, as the stop rule saw it), ghosts from `faces`, with norm
 * as above */
int mgp_residual_hilo_f64(const mgp_array* u, const mgp_array* lo, const mgp_array* f, const mgp_array* r,
                          const mgp_faces* faces, const int64_t box[6], double* norm, void* stream);

/* ------------------------------------------- FP16 storage of coarse levels
 * (no counterpart in the reference: its T is one type throughout).  Arrays
 * hold IEEE binary16; values are loaded as float, combined with the
 * consistent-mode arithmetic of the *_f32 kernels (the reference's operators:
 * gaussseidel.h:28-94, residual.h:18-81, fullweighting.h:37-122,
 * trilinearinterpolation.h:29-134) and rounded to nearest even when stored.
 * All faces Dirichlet: boxes lie strictly inside the shape (MGP_ERR_ARG
 * otherwise), colour = parity of i+j+k.
 * Scaling: the restriction into an FP16 level writes sigma * FW(r) with
 * sigma = 2^-e, where *norm_dev = max |r| = m 2^e, m in [0.5,1) (sigma = 1
 * for a zero norm); the prolongation from that level adds P(u) * 2^e with
 * the same norm.  Powers of two scale exactly. */
int mgp_fill_f16(const mgp_array* a, double value, void* stream);
/* one red (0) or black (1) half-sweep in place */
int mgp_rbgs_half_f16(const mgp_array* u, const mgp_array* f, const int64_t box[6], double omega, double h,
                      int colour, void* stream);
/* nsweeps red-then-black sweeps in place in one launch (u and f in one
 * workgroup's LDS), equal to 2 nsweeps mgp_rbgs_half_f16 calls; levels of at
 * most 8192 points (MGP_ERR_ARG otherwise) */
int mgp_rbgs_sweeps_f16(const mgp_array* u, const mgp_array* f, const int64_t box[6], double omega, double h,
                        int64_t nsweeps, void* stream);
/* r = f - L u on the box; with norm_dev, max(*norm_dev, max |r| of the
 * stored r over its shape) */
int mgp_residual_f16(const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t box[6], double h,
                     double* norm_dev, void* stream);
/* coarse = sigma * FW(fine): FP16 to FP16, and FP32 to FP16 (from the last
 * FP32 level) */
int mgp_restrict_fw_scaled_f16(const mgp_array* fine, const mgp_array* coarse, const double* norm_dev, void* stream);
int mgp_restrict_fw_scaled_f32_f16(const mgp_array* fine, const mgp_array* coarse, const double* norm_dev,
                                   void* stream);
/* fine += P(coarse) / sigma over the fine shape: FP16 to FP16, and FP16 to
 * FP32 (into the last FP32 level) */
int mgp_prolong_add_scaled_f16(const mgp_array* coarse, const mgp_array* fine, const double* norm_dev, void* stream);
int mgp_prolong_add_scaled_f16_f32(const mgp_array* coarse, const mgp_array* fine, const double* norm_dev,
                                   void* stream);

/* ------------------------------------------------------------ smoothers */
/* One red (colour 0) or black (colour 1) half-sweep in place over the box:
 * u = (1-w)u + w/6 (sum_nb u - h^2 f) (include/gaussseidel.h:28-94).  Halos
 * must hold current ghost values (see mgp_bc_neumann_*). */
int mgp_rbgs_half_f64(int mode, const mgp_array* u, const mgp_array* f, const int64_t box[6], double omega,
                      double h, int colour, void* stream);
int mgp_rbgs_half_f32(int mode, const mgp_array* u, const mgp_array* f, const int64_t box[6], double omega,
                      double h, int colour, void* stream);
/* One full red-then-black sweep in one pass, out of place: u_out = GS(u_in).
 * Neumann ghosts are formed on the fly from `faces`, as if refreshed before
 * each half-sweep (include/relaxation.h:35-52).  The shape points outside the
 * box are copied from u_in, so u_out holds what the in-place half-sweeps
 * leave in the shape; the halo of u_out is not written (ghosts are formed
 * again before every use).  Requires a
 * checkerboard colouring: always in consistent mode; in reference mode only
 * when stride[0] and stride[1] are odd (MGP_ERR_ARG otherwise).  An empty box
 * is a no-op in place and MGP_ERR_ARG out of place (nothing would be copied). */
int mgp_rbgs_sweep_f64(int mode, const mgp_array* u_in, const mgp_array* u_out, const mgp_array* f,
                       const mgp_faces* faces, const int64_t box[6], double omega, void* stream);
int mgp_rbgs_sweep_f32(int mode, const mgp_array* u_in, const mgp_array* u_out, const mgp_array* f,
                       const mgp_faces* faces, const int64_t box[6], double omega, void* stream);
/* Two such sweeps in one pass: u_out = GS(GS(u_in)), bit for bit the result of
 * mgp_rbgs_sweep_* twice through a second buffer (the GaussSeidel smoother's
 * pairs of sweeps, include/gaussseidel.h:28-94 run twice per point; no
 * single counterpart in the reference), shape points outside the box copied
 * as above.  u_out must be a different buffer; MGP_ERR_ARG also where the
 * level has fewer than
 * 3 x-planes or a plane of 1 GiB or more (run single sweeps there), and for an
 * empty box. */
int mgp_rbgs_sweep2_f64(int mode, const mgp_array* u_in, const mgp_array* u_out, const mgp_array* f,
                        const mgp_faces* faces, const int64_t box[6], double omega, void* stream);
int mgp_rbgs_sweep2_f32(int mode, const mgp_array* u_in, const mgp_array* u_out, const mgp_array* f,
                        const mgp_faces* faces, const int64_t box[6], double omega, void* stream);
/* The cycle's correction and the two sweeps after it in one pass: u_out =
 * GS(GS(u_in + P(coarse))) on the box and u_in + P(coarse) at the shape
 * points outside it, bit for bit mgp_prolong_add_* of coarse into u_in
 * followed by mgp_rbgs_sweep2_* (the prolongation and correction of
 * include/vcycle.h:81-94 and the first two post-smoothing sweeps; no single
 * counterpart in the reference).  u_in is not changed.  Consistent mode only;
 * coarse 2:1 nested in u_in (shape 2(n-1)+1 per axis) and neither u_in nor
 * u_out; otherwise, and where mgp_rbgs_sweep2_* refuses the level,
 * MGP_ERR_ARG. */
int mgp_rbgs_sweep2_prolong_f64(int mode, const mgp_array* u_in, const mgp_array* u_out, const mgp_array* f,
                                const mgp_faces* faces, const int64_t box[6], double omega,
                                const mgp_array* coarse, void* stream);
int mgp_rbgs_sweep2_prolong_f32(int mode, const mgp_array* u_in, const mgp_array* u_out, const mgp_array* f,
                                const mgp_faces* faces, const int64_t box[6], double omega,
                                const mgp_array* coarse, void* stream);
/* The two sweeps of u = 0: u_out = GS(GS(0)) on the box and 0 at the shape
 * points outside it, bit for bit mgp_fill_* of 0 into a buffer followed by
 * mgp_rbgs_sweep2_* from it, without the fill and without reading u (the
 * first smoothing of a cycle for the correction of iterative refinement,
 * whose initial guess is zero; no counterpart in the reference).  u_out's halo
 * is not written.  Consistent mode only; MGP_ERR_ARG as for mgp_rbgs_sweep2_*. */
int mgp_rbgs_sweep2_from_zero_f64(int mode, const mgp_array* u_out, const mgp_array* f, const mgp_faces* faces,
                                  const int64_t box[6], double omega, void* stream);
int mgp_rbgs_sweep2_from_zero_f32(int mode, const mgp_array* u_out, const mgp_array* f, const mgp_faces* faces,
                                  const int64_t box[6], double omega, void* stream);
/* nsweeps such sweeps in one launch, for a level small enough for one
 * workgroup's LDS ((nx+2)(ny+2)(nz+2) + nx ny nz elements <= 64 KiB;
 * MGP_ERR_ARG otherwise): the result of the last sweep goes to u_out, which may
 * be u_in.  Same values as nsweeps calls of mgp_rbgs_sweep_* (the coarsest
 * level's smoothing, where launch latency dominates).  An empty box or
 * nsweeps = 0 is a no-op in place and MGP_ERR_ARG out of place. */
int mgp_rbgs_sweeps_f64(int mode, const mgp_array* u_in, const mgp_array* u_out, const mgp_array* f,
                        const mgp_faces* faces, const int64_t box[6], double omega, int64_t nsweeps, void* stream);
int mgp_rbgs_sweeps_f32(int mode, const mgp_array* u_in, const mgp_array* u_out, const mgp_array* f,
                        const mgp_faces* faces, const int64_t box[6], double omega, int64_t nsweeps, void* stream);
/* The same on FP32 arrays with FP64 arithmetic: u and f widened (exactly) to
 * double in LDS, the nsweeps sweeps in FP64, the result rounded once to FP32
 * (the coarsest-level solve of FP32 cycles inside FP64 refinement; no
 * counterpart in the reference).  The level must fit the LDS limit above with
 * 8-byte elements. */
int mgp_rbgs_sweeps_f32_f64(int mode, const mgp_array* u_in, const mgp_array* u_out, const mgp_array* f,
                            const mgp_faces* faces, const int64_t box[6], double omega, int64_t nsweeps, void* stream);
/* The coarsest level solved by conjugate gradients in FP64 in one workgroup
 * (consistent mode; no counterpart in the reference, whose coarsest level
 * gets the smoothing passes of the others, include/vcycle.h:27-31): with
 * res = f - L u (Neumann ghosts sign*2h*g + mirror from `faces`, Dirichlet
 * boundary nodes as stored), CG on the symmetric form -W L0 d = -W res of the
 * homogeneous operator L0 (W = 1/2 per Neumann boundary coordinate) from d = 0
 * until |r| <= rtol |r0| or max_iters steps, then u += d in place.  The box
 * must be the level's natural box (boundary nodes of Dirichlet faces out,
 * those of Neumann faces in) and the level small enough for one workgroup:
 * (nx+2)(ny+2)(nz+2) doubles in 64 KiB LDS, at most 8192 box points and 4096
 * points per face (MGP_ERR_ARG otherwise).  At least one face Dirichlet
 * (else the system is singular).  Fixed-order reductions: the same bits on
 * every run. */
int mgp_cg_solve_f64(const mgp_array* u, const mgp_array* f, const mgp_faces* faces, const int64_t box[6],
                     int64_t max_iters, double rtol, void* stream);
int mgp_cg_solve_f32(const mgp_array* u, const mgp_array* f, const mgp_faces* faces, const int64_t box[6],
                     int64_t max_iters, double rtol, void* stream);
/* Weighted Jacobi over the box: u = (1-w)v + w/6 (sum_nb v - h^2 f)
 * (include/jacobi.h:39-82).  v's halos must hold current ghost values. */
int mgp_jacobi_f64(int mode, const mgp_array* u, const mgp_array* v, const mgp_array* f, const int64_t box[6],
                   double omega, double h, void* stream);
int mgp_jacobi_f32(int mode, const mgp_array* u, const mgp_array* v, const mgp_array* f, const int64_t box[6],
                   double omega, double h, void* stream);

/* ------------------------------------------------------------- residual */
/* r = f - L_h u over the box (include/residual.h:58-79); u's halos must hold
 * current ghost values.  If norm_dev != NULL, also *norm_dev = max(*norm_dev,
 * max |r| over r's whole shape) in the same pass. */
int mgp_residual_f64(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t box[6],
                     double h, double* norm_dev, void* stream);
int mgp_residual_f32(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t box[6],
                     double h, double* norm_dev, void* stream);

/* ----------------------------------------------------- grid transfers */
/* out(i,j,k) = in(2i,2j,2k) over out's shape (include/injection.h:29-67) */
int mgp_restrict_inject_f64(const mgp_array* in, const mgp_array* out, void* stream);
int mgp_restrict_inject_f32(const mgp_array* in, const mgp_array* out, void* stream);
/* 27-point (1,2,1)^3/64 full weighting with neighbour indices clamped to the
 * shape (include/fullweighting.h:37-122); also valid on face arrays */
int mgp_restrict_fw_f64(const mgp_array* in, const mgp_array* out, void* stream);
int mgp_restrict_fw_f32(const mgp_array* in, const mgp_array* out, void* stream);
/* fine = trilinear interpolation of coarse over fine's whole shape
 * (include/trilinearinterpolation.h:29-134), bit-identical to its four passes */
int mgp_prolong_trilinear_f64(const mgp_array* coarse, const mgp_array* fine, void* stream);
int mgp_prolong_trilinear_f32(const mgp_array* coarse, const mgp_array* fine, void* stream);
/* u += trilinear interpolation of coarse, over u's whole shape, in one pass:
 * the reference's prolong-into-scratch followed by DeviceArray::add
 * (include/vcycle.h:81-94) without the scratch round trip */
int mgp_prolong_add_f64(const mgp_array* coarse, const mgp_array* u, void* stream);
int mgp_prolong_add_f32(const mgp_array* coarse, const mgp_array* u, void* stream);

/* ------------------------------------------------- boundary conditions */
/* Dirichlet: boundary nodes of face `loc` = g (include/dirichlet.h:50-93) */
int mgp_bc_dirichlet_f64(const mgp_array* u, int loc, const void* g, void* stream);
int mgp_bc_dirichlet_f32(const mgp_array* u, int loc, const void* g, void* stream);
/* Neumann: ghost layer of face `loc` = sign*2h*g + u(mirror), sign -1 on
 * WEST/SOUTH/BOTTOM (include/neumann.h:63-131); g NULL = homogeneous */
int mgp_bc_neumann_f64(const mgp_array* u, int loc, const void* g, double h, void* stream);
int mgp_bc_neumann_f32(const mgp_array* u, int loc, const void* g, double h, void* stream);
/* The ghost layers of every face of type MGP_NEUMANN in `faces` in one launch
 * (faces->g[q], NULL = homogeneous; faces->h): the boundary refresh before a
 * residual (include/residual.h:26-34, include/neumann.h:133-144 per face).  The
 * faces write disjoint ghost points and read inside the shape, so this equals
 * mgp_bc_neumann_* on each Neumann face in any order. */
int mgp_bc_neumann_faces_f64(const mgp_array* u, const mgp_faces* faces, void* stream);
int mgp_bc_neumann_faces_f32(const mgp_array* u, const mgp_faces* faces, void* stream);
/* Coarse Neumann data as the reference forms it (include/neumann.h:146-252):
 * g_coarse = R(g_fine) - (c1 u(b) + c2 u(b+-1) + c3 u(b+-2))/h_fine with
 * c = (-3/2, 2, -1/2) on WEST/SOUTH/BOTTOM and (3/2, -2, 1/2) elsewhere;
 * restriction = MGP_INJECTION or MGP_FULL_WEIGHTING */
int mgp_bc_restrict_neumann_f64(int restriction, const mgp_array* u_fine, int loc, const void* g_fine,
                                void* g_coarse, double h_fine, void* stream);
int mgp_bc_restrict_neumann_f32(int restriction, const mgp_array* u_fine, int loc, const void* g_fine,
                                void* g_coarse, double h_fine, void* stream);

/* -------------------------------------------------------- graphs, timing */
/* Record the work enqueued on `stream` between begin and end into a graph
 * (hipStreamBeginCapture, relaxed mode); *exec receives an instantiated graph */
int mgp_graph_begin(void* stream);
int mgp_graph_end(void* stream, void** exec);
int mgp_graph_launch(void* exec, void* stream);
int mgp_graph_destroy(void* exec);
int mgp_event_create(void** event);
int mgp_event_destroy(void* event);
int mgp_event_record(void* event, void* stream);
int mgp_event_sync(void* event);
int mgp_event_elapsed_ms(void* start, void* stop, float* ms);

/* --------------------------------------------------------- whole solver
 * A PoissonSolver<T, R, TrilinearInterpolation, S> (include/libpoisson.h:28-67
 * of the reference) behind an opaque handle: the calls an FFI binding of the
 * reference's solver façade makes.  Host arrays returned by
 * mgp_solver_host_array are the solver's own mirrors (Array<T>::at): write
 * them before mgp_solver_to_device, read them after mgp_solver_to_host. */
enum { MGP_JACOBI = 0, MGP_GAUSS_SEIDEL = 1 };
#define MGP_SOLVER_NO_GRAPHS 1 /* run cycles launch by launch (Settings::use_graphs = false) */
#define MGP_SOLVER_NO_FUSED 2  /* red-black half-sweeps only (Settings::fused_sweeps = false) */
#define MGP_SOLVER_QUIET 4     /* no per-cycle timing line, no level listing at creation
                                  (Settings::print_cycle_time, print_settings = false) */
/* MGP_MIXED: the iterate as one FP64 array (no FP32 lo word; the FP64
 * residual then floors near 12/h^2 ulp(u)) (Settings::refine_hilo = false) */
#define MGP_SOLVER_PLAIN_ITERATE 8
/* MGP_MIXED: the FP32 cycle's coarsest level smoothed in FP32 arithmetic
 * instead of FP64 (Settings::refine_coarse_fp64 = false) */
#define MGP_SOLVER_COARSE_FP32 16
/* consistent mode: every visit of the coarsest level solves it by CG in FP64
 * (mgp_cg_solve_*, to 1e-8 relative) instead of coarse_sweeps smoothing passes,
 * where the level fits one workgroup (Settings::coarse_cg) */
#define MGP_SOLVER_COARSE_CG 32

typedef struct mgp_solver_config {
    int64_t dims[3];     /* Settings::dims (include/settings.h:17) */
    int64_t levels;      /* Settings::levels */
    double h;            /* Settings::h */
    double origin[3];    /* Settings::origin */
    int32_t bc[6];       /* BoundaryCondition, Location_t order N,S,E,W,T,B: MGP_DIRICHLET / MGP_NEUMANN */
    int32_t dtype;       /* T: MGP_F64 / MGP_F32, or MGP_MIXED: RefinementSolver (FP32 cycles
                            inside FP64 iterative refinement; needs MGP_MODE_CONSISTENT) */
    int32_t restriction; /* R: MGP_INJECTION / MGP_FULL_WEIGHTING */
    int32_t relaxation;  /* S: MGP_JACOBI / MGP_GAUSS_SEIDEL */
    int32_t mode;        /* MGP_MODE_REFERENCE / MGP_MODE_CONSISTENT */
    int32_t flags;       /* MGP_SOLVER_* */
    int64_t maxiter;     /* Settings::maxiter */
    int64_t miniter;     /* Settings::miniter */
    double tolerance;    /* Settings::tolerance */
    int64_t coarse_sweeps; /* Settings::coarse_sweeps (consistent mode; 0 = nsmooth) */
    int64_t refine_polish; /* MGP_MIXED: FP64 smoothing passes on the finest level after each
                              update (Settings::refine_polish; 0 = none) */
    int64_t half_from;     /* MGP_MIXED: levels >= half_from of the FP32 cycle stored in FP16
                              (Settings::half_from; 0 = none; needs 2 <= half_from < levels,
                              Gauss-Seidel, full weighting, all faces Dirichlet) */
} mgp_solver_config;

/* PoissonSolver(settings, BC) (libpoisson.h:79-87) */
int mgp_solver_create(const mgp_solver_config* cfg, void** solver);
int mgp_solver_destroy(void* solver);
/* init() with a built-in analytic problem (0: sin(2 pi x)sin(2 pi y)sin(2 pi z);
 * 1-3: the reference's PROBLEM1-3, src/problem_definition.cpp) (libpoisson.h:159-171) */
int mgp_solver_init_problem(void* solver, int problem);
/* init() with caller functions u, f, du/dx, du/dy, du/dz */
int mgp_solver_init_fn(void* solver, double (*u)(double, double, double), double (*f)(double, double, double),
                       double (*dudx)(double, double, double), double (*dudy)(double, double, double),
                       double (*dudz)(double, double, double));
int mgp_solver_init_zero(void* solver);  /* libpoisson.h:174-190 */
int mgp_solver_to_device(void* solver);  /* libpoisson.h:193-209 */
/* the device state to_device() gives right after creation with every array
 * zero (init_zero), made on the device without a host transfer (no
 * counterpart in the reference) */
int mgp_solver_to_device_zero(void* solver);
int mgp_solver_to_host(void* solver);    /* libpoisson.h:212-228 */
/* solve(solver, nsmooth, omega, "", maxtime) (libpoisson.h:237-357);
 * solver = "Vcycle"/"vcycle"/"Fcycle"/"fcycle"/"Relaxation"/"relaxation" */
int mgp_solver_solve(void* solver, const char* cycle, int64_t nsmooth, double omega, double maxtime);
/* one cycle without norms, enqueued; then mgp_solver_relres for ||r||/||f|| */
int mgp_solver_cycle(void* solver, const char* cycle, int64_t nsmooth, double omega);
int mgp_solver_relres(void* solver, double* relres);
int mgp_solver_sync(void* solver);
/* every level's u back to zero with its boundary data, on the device, for a
 * fresh solve with unchanged right-hand side and faces (no host transfer) */
int mgp_solver_reset(void* solver);
/* relative_residual(), solve_time(), solve_iterations() (libpoisson.h:378-393) */
int mgp_solver_stats(void* solver, double* rel_res, double* seconds, int64_t* iterations);
/* host mirror and device descriptor of level `level`'s u ('u'), uprev ('p'),
 * f ('f') or r ('r'); for MGP_MIXED also 'l': the FP32 lo word of the hi + lo
 * iterate on level 0 (device only, *host = NULL; MGP_ERR_ARG with
 * MGP_SOLVER_PLAIN_ITERATE) */
int mgp_solver_array(void* solver, int level, int which, void** host, mgp_array* device);
/* host mirror and device descriptor of a face array; *type receives the
 * effective boundary type (MGP_DIRICHLET / MGP_NEUMANN).  The face arrays of
 * levels >= 1 belong to the solver: each descent sets them (zero in consistent
 * mode, where they are filled once and kept), so a caller writes level 0's */
int mgp_solver_face(void* solver, int level, int loc, void** host, mgp_array* device, int* type);
/* VTK of level 0's u (save, libpoisson.h:359-363) */
int mgp_solver_save(void* solver, const char* path);
/* message of the last failing mgp_solver_* / mgp_batch_* call on this thread */
const char* mgp_last_error(void);

/* max |f - L u| over the box, and r's stored values outside it, folded into
 * *norm (a device double the caller zeroed); r is not written.  The
 * convergence check of a solve, which needs the norm only (the solve writes
 * r once, for the final u).  MGP_ERR_ARG when the layouts do not allow the
 * one-pass form. */
int mgp_residual_norm_f64(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t box[6],
                          double h, double* norm, void* stream);
int mgp_residual_norm_f32(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t box[6],
                          double h, double* norm, void* stream);

/* Consistent mode: coarse f = full weighting (include/fullweighting.h:59-121)
 * of r = f - L u (include/residual.h:58-79), r computed on the box and not
 * stored, r outside the box read as stored; u's halos must hold current
 * ghosts.  The descent step of a cycle (residual, then restriction) in one
 * pass over the fine level.  MGP_ERR_ARG in reference mode or when the
 * layouts do not allow it. */
int mgp_residual_restrict_fw_f64(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r,
                                 const mgp_array* coarse_f, const int64_t box[6], double h, void* stream);
int mgp_residual_restrict_fw_f32(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r,
                                 const mgp_array* coarse_f, const int64_t box[6], double h, void* stream);

/* residual of FP64 u and f on the box, written as FP32 into `out` (same
 * shape, halos and element strides as u) instead of r; with norm, max |r|
 * over the box and r's stored values outside it.  r itself is only read
 * (outside the box).  The residual and cast of mixed-precision refinement in
 * one pass (no counterpart in the reference). */
int mgp_residual_cast_f64_f32(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r,
                              const mgp_array* out, const int64_t box[6], double h, double* norm, void* stream);

/* the shape of a (halos dropped) into the dense C-order device buffer `out`
 * (shape[0]*shape[1]*shape[2] doubles); a result copy needs one transfer then
 * (no counterpart in the reference) */
int mgp_pack_f64(const mgp_array* a, void* out, void* stream);

/* synthetic right-hand side over the shape of f: sum over nmodes (<= 16) of
 * amp[m] sin(a pi x) sin(b pi y) sin(c pi z), (a, b, c) = abc[3m..3m+2], at
 * x = i h, y = j h, z = k h.  No counterpart in the reference (its f comes
 * from compile-time analytic functions). */
#define MGP_MAX_MODES 16
int mgp_fill_modes_f64(const mgp_array* f, int64_t nmodes, const int32_t* abc, const double* amp, double h,
                       void* stream);

/* ---------------------------------------------------------------- batched solves
 * Many independent problems of uneven shapes solved together: zero Dirichlet
 * data on every face, h = 1/(nx-1), FP32 V-cycles (full weighting, trilinear
 * prolongation, the chosen smoother) inside FP64 iterative refinement on
 * ||r||inf/||f||inf of the FP64 residual, one launch per operation and level
 * for all problems, a converged problem drops out of later launches.  Each
 * problem's result is bit-identical to a single solver created with
 * dtype MGP_MIXED, mode MGP_MODE_CONSISTENT, bc all MGP_DIRICHLET, restriction
 * MGP_FULL_WEIGHTING and the same levels/smoother/tolerances, solved with
 * "Vcycle".  No counterpart in the reference (one problem per PoissonSolver).
 * Tile shapes per level size are chosen by timing candidates at creation and
 * cached per process. */
#define MGP_BATCH_NO_GRAPHS 1 /* launch kernels one by one (no recorded iteration) */
#define MGP_BATCH_RETUNE 2    /* time the tile candidates even when cached */
typedef struct {
    int64_t count;         /* problems */
    int64_t levels;        /* levels per problem; 0: down to a 3-point coarsest axis */
    int32_t relaxation;    /* MGP_JACOBI or MGP_GAUSS_SEIDEL */
    int32_t flags;         /* MGP_BATCH_* */
    int64_t nsmooth;       /* pre- and post-smoothing passes */
    double omega;          /* <= 0: 6/7 (Jacobi) or 8/9 (Gauss-Seidel), as the reference's defaults */
    int64_t coarse_sweeps; /* passes on a problem's coarsest level; 0: nsmooth */
    int64_t maxiter, miniter;
    double tolerance;
} mgp_batch_config;
/* dims: count x 3 axis sizes (each 2^k m + 1 with the levels' coarsening) */
int mgp_batch_create(const mgp_batch_config* config, const int64_t* dims, void** batch);
int mgp_batch_destroy(void* batch);
/* host f of problem p over its whole shape, C order (x slowest) */
int mgp_batch_set_rhs(void* batch, int64_t p, const double* f);
/* solve every problem from u = 0; *seconds (optional) = wall time */
int mgp_batch_solve(void* batch, double* seconds);
int mgp_batch_stats(void* batch, int64_t p, double* rel_res, int64_t* iterations);
/* host copy of problem p's FP64 solution over its shape */
int mgp_batch_solution(void* batch, int64_t p, double* out);
/* tile chosen by the autotuner for problem p's level: planes, rows per item */
int mgp_batch_tile(void* batch, int64_t p, int64_t level, int64_t out[2]);

/* ---------------------------------------------------------------- serving
 * A request loop on one GPU: solves of mixed shapes and face types, each with
 * zero boundary data and a Fourier-mode right-hand side (mgp_fill_modes_f64),
 * accepted as they arrive (mgp_server_submit returns at once) and run by a
 * worker thread as FP32 F-cycles inside FP64 iterative refinement (consistent
 * mode, full weighting, levels down to a 3-point coarsest axis).  Requests are
 * admitted in arrival order while fewer than max_active run and their
 * hierarchy fits in hbm_budget next to the active and resident ones;
 * hierarchies stay resident for later requests of the same shape and face
 * types (least recently used evicted first).  max_active worker threads
 * each run one solve at a time on a stream of their own, so the solves' work
 * overlaps on the device.  Hierarchies are carved from one device region of
 * hbm_budget * 9/8 bytes (within free memory) allocated at creation, so new
 * shapes and evictions make no hipMalloc / hipFree (an array that finds no
 * free range large enough in the region is allocated on its own).  A result is the same bits
 * whether its request ran alone or among others.  All-Neumann requests are
 * rejected (singular).  No counterpart in the reference. */
typedef struct {
    int64_t hbm_budget;     /* bytes of device memory for active + resident hierarchies */
    int32_t max_active;     /* solves in flight at once (worker threads, one stream each) */
    int32_t relaxation;     /* MGP_GAUSS_SEIDEL or MGP_JACOBI */
    int64_t nsmooth;        /* smoothing passes (pre = post) */
    double omega;           /* <= 0: the smoother's default */
    int64_t coarse_sweeps;  /* passes on a coarsest level too large for the CG solve
                               (mgp_cg_solve_*, used wherever the level fits); 0: nsmooth */
    int64_t maxiter;        /* refinement iterations at most */
    double tolerance;       /* on ||r||inf/||f||inf of the FP64 residual */
    int32_t flags;          /* MGP_SERVER_* */
} mgp_server_config;
/* record each hierarchy's cycles as HIP graphs (replayed by later iterations
 * and requests of the same shape); without it every kernel is launched
 * directly, which is cheaper for hierarchies used for a few iterations only */
#define MGP_SERVER_GRAPHS 1
/* allocate each hierarchy array by array with hipMalloc (and free it with
 * hipFree on eviction) instead of from the server's device region */
#define MGP_SERVER_DIRECT_ALLOC 2
typedef struct {
    int64_t dims[3];                   /* 2^k m + 1 per axis; h = 1/(dims[0]-1) */
    int32_t bc[6];                     /* MGP_DIRICHLET / MGP_NEUMANN, order N S E W T B */
    int32_t nmodes;                    /* right-hand side: nmodes (<= 16) Fourier modes */
    int32_t abc[3 * MGP_MAX_MODES];
    double amp[MGP_MAX_MODES];
    double* u_out;                     /* host, dims[0]*dims[1]*dims[2] doubles (C order), or NULL */
} mgp_request;
typedef struct {
    int32_t status;                    /* MGP_OK or an error code */
    int64_t iterations;
    double rel_res;
    double latency_s;                  /* submit -> result on the host */
    double queue_s;                    /* submit -> admission */
    double run_s;                      /* admission -> result on the host */
    double setup_s;                    /* admission: slot (hierarchy) creation or reuse, f, u = 0 */
    double fetch_s;                    /* solution copy to the host */
} mgp_result;
int mgp_server_create(const mgp_server_config* config, void** server);
/* waits for every submitted request to finish */
int mgp_server_destroy(void* server);
int mgp_server_submit(void* server, const mgp_request* request, int64_t* id);
/* blocks until request `id` is done; each id can be waited for once */
int mgp_server_wait(void* server, int64_t id, mgp_result* result);
int mgp_server_stats(void* server, int64_t* completed, int64_t* peak_active, int64_t* bytes_resident);

/* ------------------------------------------------ cycles without allocation
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ncycles V- or F-cycles ("Vcycle"/"vcycle"/"Fcycle"/
 * "fcycle") on the caller's finest-level u and f, the values that ncycles
 * mgp_solver_cycle calls give on a solver created with dtype MGP_F64, mode
 * MGP_MODE_CONSISTENT, MGP_GAUSS_SEIDEL, MGP_FULL_WEIGHTING and the same dims,
 * levels, coarse_sweeps, u, f and faces (omega < 0: 8/9, as that solver).  u
 * and f have that solver's level-0 layout (one halo layer on each Neumann face,
 * none on Dirichlet faces, dense strides); faces gives the face types, the
 * Neumann data (required) and h of level 0.  The finest level's residual and
 * second u buffer and every coarser level live in one library-owned device
 * array of MGP_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: the
 * call allocates nothing and does not synchronise.  With norm != NULL, *norm
 * (a device double) = max |f - L u| over the box of the final u (written, not
 * accumulated).  MGP_ERR_ARG when the hierarchy does not fit the array (level
 * 0 of about 2^21 points or more) or the arguments do not fit the above.  Every
 * call clears the parts of the array it reads before use, but calls that
 * overlap on one GPU share the array and race. */
#define MGP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int mgp_cycle_noalloc_f64(const char* cycle, const mgp_array* u, const mgp_array* f, const mgp_faces* faces,
                          int64_t levels, int64_t ncycles, int64_t nsmooth, double omega, int64_t coarse_sweeps,
                          double* norm, void* stream);

/* ---------------------------------------------------------------- tuning
 * Launch-shape knobs (process-wide; take effect for the next launches and for
 * graphs recorded afterwards).  Keys: "sweep.kernel" (1: workgroup tiles
 * sharing planes through LDS; 2: one wave per tile, values in registers,
 * default), "sweep.target_wgs" (workgroups, or waves for kernel 2, a one-pass
 * sweep aims for; default 16384), "sweep.min_chunk" (shortest x march per
 * tile, default 4), "sweep.waves" (4, 8 or 16 waves per workgroup, kernel 1,
 * default 8), "sweep.rows" (8, 12 or 16 register rows per wave tile, kernel 2;
 * 0 = 12 for FP64 and 16 for FP32, default), "sweep.depth" (1 or 2 planes of
 * loads in flight per wave, kernel 2, default 1), "sweep.pair" (1: the
 * GaussSeidel smoother runs its sweeps in pairs through mgp_rbgs_sweep2_*,
 * default; 0: one pass per sweep), "sweep.pair_min_points" (levels with fewer
 * points run one pass per sweep; default 2^25), "sweep.small_points" and
 * "sweep.small_min_chunk" (levels with fewer points than the first use the
 * second as their shortest x march; defaults 2^18 and 1), "sweep2.rows"
 * (register rows per two-sweep
 * wave tile: 12 or 16 for FP64, 16, 20 or 24 for FP32; 0 = 16, default),
 * "sweep2.chunks_per_tile" (least x chunks per two-sweep wave tile, made odd,
 * and at least enough for "sweep2.target_waves" waves; defaults 15 and 6800), "sweep2.prolong" (2:
 * in consistent mode the cycles add the prolongation of the coarse correction
 * inside the first two-sweep pass after it, mgp_rbgs_sweep2_prolong_*, on FP64
 * and FP32 levels, default; 1: on FP64 levels only; 0: mgp_prolong_add_* and
 * then the sweeps), "sweep2.pack" (consistent-mode two-sweep stages evaluate each pair
 * of register rows once, each lane taking the row of the stage's colour: bit 0
 * FP64, bit 1 FP32; default 3), "residual.target_wgs",
 * "residual_fw.target_wgs", "prolong.target_wgs" (waves the norm-only
 * residual, the fused residual + full-weighting and the prolongation marches
 * aim for; 0 = the value of
 * "sweep.target_wgs"; defaults 4096, 32768, 65536), "update_residual.target_wgs"
 * (waves of mgp_update_residual_f64_f32 and the hi + lo kernels; default 65536),
 * "update_residual.rows" (8 or 16 register rows per wave tile of the hi + lo
 * kernels, default 16), "residual.kernel" (1:
 * per-point kernels; 2: one wave per tile marching along x, default),
 * "residual.rows" (8 or 16 register rows per residual tile, default 16),
 * "residual_fw.rows" (4 or 6 coarse rows per wave tile of the fused residual
 * and full-weighting restriction; 0 = 6, default), "residual_fw.depth" (1 or
 * 2 fine planes of loads in flight per wave; 0 = 1 for FP64 and 2 for FP32,
 * default),
 * "prolong.kernel" (1: per-point kernel; 2: wave tiles marching along x for
 * mgp_prolong_add_*, default), "batch.blocks" (workgroups per launch of the
 * batched solver's work lists, default 16384).
 * No counterpart in the
 * reference, whose launch
 * shape was fixed by the OpenMP runtime. */
int mgp_tuning_set(const char* key, int64_t value);
int mgp_tuning_get(const char* key, int64_t* value);

/* ---------------------------------------------------------------- info */
const char* mgp_error_string(int code);
const char* mgp_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MGP_H */
