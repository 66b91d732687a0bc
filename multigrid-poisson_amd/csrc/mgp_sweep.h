// One-pass red-black sweep (mgp_sweep.hip), called by mgp_rbgs_sweep_*.
#pragma once

#include <stdint.h>

#include "mgp.h"

namespace mgp {
// Arguments validated by the caller (checkerboard colouring, box inside the
// shape, Neumann faces wherever the box touches the boundary).
template <class T>
int march_sweep(int mode, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* faces,
                const int64_t box[6], double omega, int64_t xchunk, void* stream);
// two sweeps in one pass (ui -> uo, ui != uo); MGP_ERR_ARG when the shape
// does not allow it (fewer than 3 planes, or a plane of 1 GiB or more).  With
// coarse (2:1 nested, checked by the caller; consistent mode only): the two
// sweeps of ui + P(coarse), the points outside the box receiving ui +
// P(coarse).  With zero_in (consistent mode only): the two sweeps of u = 0, ui
// giving the layout only (not read), the points outside the box receiving 0
template <class T>
int march_sweep2(int mode, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* faces,
                 const int64_t box[6], double omega, void* stream, const mgp_array* coarse = nullptr,
                 bool zero_in = false);
extern int64_t g_sweep_pair;
extern int64_t g_sweep2_rows;
extern int64_t g_sweep2_prolong;
extern int64_t g_sweep2_pack;
extern int64_t g_sweep2_waves;
// residual on a box through wave tiles (norm: max |r| over the whole shape
// folded into *norm); MGP_ERR_ARG when the layouts do not allow it
template <class T>
int march_residual(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t box[6],
                   double h, double* norm, void* stream);
// the same with r written, in type O, to `out` (same element layout) instead of r
// (store = false: only the norm; r and out untouched)
template <class T, class O>
int march_residual_to(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r, const mgp_array* out,
                      const int64_t box[6], double h, double* norm, void* stream, bool store = true);
// u_new = u + (double) e over the shape, then (float)(f - L u_new) into out on
// the box with Neumann ghosts formed from u_new (max |r| into *norm, r's
// stored values outside the box included); u_new a different buffer
int march_update_residual(int mode, const mgp_array* u, const mgp_array* un, const mgp_array* e, const mgp_array* f,
                          const mgp_array* r, const mgp_array* out, const mgp_faces* faces, const int64_t box[6],
                          double* norm, void* stream);
// iterate held as hi + lo (consistent mode): with e, (un, ln) = hi + lo + e
// renormalised and (float)(f - L(un + ln)) into out; without e (e, un, ln
// NULL), (double)(f - L(u + lo)) into out.  Ghosts formed from faces, max |r|
// into *norm (r's stored values outside the box included)
int march_hilo_residual(const mgp_array* u, const mgp_array* un, const mgp_array* lo, const mgp_array* ln,
                        const mgp_array* e, const mgp_array* f, const mgp_array* r, const mgp_array* out,
                        const mgp_faces* faces, const int64_t box[6], double* norm, void* stream);
extern int64_t g_hilo_rows;
// u += trilinear P(coarse) over the fine shape through wave tiles
template <class T>
int march_prolong_add(const mgp_array* coarse, const mgp_array* fine, void* stream);
extern int64_t g_prolong_kernel;
// coarse f = full weighting of (f - L u) on the fine box in one pass
// (consistent mode; r outside the box read as stored, r itself not written)
template <class T>
int march_residual_fw(const mgp_array* u, const mgp_array* f, const mgp_array* r, const mgp_array* cf,
                      const int64_t box[6], double h, void* stream);
extern int64_t g_fw_rows;
extern int64_t g_fw_depth;
extern int64_t g_resid_kernel;
extern int64_t g_resid_rows;
// default x-chunk length for a box
int64_t march_xchunk(const int64_t box[6]);
extern int64_t g_batch_blocks;  // workgroups per batched launch (mgp_batch.hip)
extern int64_t g_sweep_target_wgs;
extern int64_t g_sweep2_cpt;
extern int64_t g_resid_target, g_fw_target, g_prolong_target, g_upd_target;
extern int64_t g_small_points, g_small_min_chunk, g_pair_min_points;
extern int64_t g_sweep_min_chunk;
extern int64_t g_sweep_waves;
extern int64_t g_sweep_kernel;
extern int64_t g_sweep_rows;
extern int64_t g_sweep_depth;
}  // namespace mgp
