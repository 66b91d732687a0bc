// gfx950 kernels of the mgp library and their C entry points.
//
// Each kernel restates one OpenMP target region of the reference
// (eljens/multigrid-poisson, include/*.h) as a HIP kernel; the cited
// file:line is the region it replaces.  Reference-mode expressions keep the
// reference's operand order and C++ types term by term (this file is built
// with -ffp-contract=off), so results are bit-identical to the reference's
// host build.  Consistent-mode expressions are this library's own
// cancellation-aware forms with explicit fma.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mgp.h"
#include "mgp_device.h"
#include "mgp_point.h"
#include "mgp_sweep.h"

namespace mgp {

struct Box {
    int64_t x0, x1, y0, y1, z0, z1;
};

static inline Box box_of(const int64_t b[6]) { return Box{b[0], b[1], b[2], b[3], b[4], b[5]}; }
static inline bool box_empty(const Box& b) { return b.x1 <= b.x0 || b.y1 <= b.y0 || b.z1 <= b.z0; }

// Launch geometry over a box: 64 threads along z (contiguous), 4 rows of y;
// grid.x enumerates (x, z-chunk) pairs, grid.y the y-row groups.
constexpr int BZ = 64;
constexpr int BY = 4;

struct BoxGrid {
    dim3 grid, block;
    int64_t zchunks;
};

static inline BoxGrid box_grid(const Box& b) {
    BoxGrid g;
    g.zchunks = (b.z1 - b.z0 + BZ - 1) / BZ;
    g.grid = dim3((unsigned)((b.x1 - b.x0) * g.zchunks), (unsigned)((b.y1 - b.y0 + BY - 1) / BY), 1);
    g.block = dim3(BZ, BY, 1);
    return g;
}

// (i, j, k) of this thread for a BoxGrid launch; false when outside the box.
__device__ __forceinline__ bool box_point(const Box& b, int64_t zchunks, int64_t& i, int64_t& j, int64_t& k) {
    const int64_t bx = blockIdx.x;
    i = b.x0 + bx / zchunks;
    k = b.z0 + (bx % zchunks) * BZ + threadIdx.x;
    j = b.y0 + (int64_t)blockIdx.y * BY + threadIdx.y;
    return k < b.z1 && j < b.y1;
}

// Column geometry for reductions: one block per (z-chunk, y-group) column,
// marching over every x plane, so a whole reduction issues one atomic per
// column instead of one per plane.
static inline BoxGrid column_grid(const Box& b) {
    BoxGrid g;
    g.zchunks = (b.z1 - b.z0 + BZ - 1) / BZ;
    g.grid = dim3((unsigned)g.zchunks, (unsigned)((b.y1 - b.y0 + BY - 1) / BY), 1);
    g.block = dim3(BZ, BY, 1);
    return g;
}

__device__ __forceinline__ bool column_point(const Box& b, int64_t& j, int64_t& k) {
    k = b.z0 + (int64_t)blockIdx.x * BZ + threadIdx.x;
    j = b.y0 + (int64_t)blockIdx.y * BY + threadIdx.y;
    return k < b.z1 && j < b.y1;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Max over the block of non-negative v, folded into *out with a 64-bit atomic
// max on the IEEE bit pattern (order-independent, so the result is exact).
__device__ __forceinline__ void block_max_into(double v, unsigned long long* out) {
    __shared__ double part[16];
    const int lane = (threadIdx.y * blockDim.x + threadIdx.x) & 63;
    const int wave = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (blockDim.x * blockDim.y + 63) >> 6;
    v = wave_max(v);
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (wave == 0) {
        double w = lane < nwaves ? part[lane] : 0.0;
        w = wave_max(w);
        if (lane == 0 && w > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(w));
    }
}

// ------------------------------------------------------------ elementwise

template <class T>
__global__ void k_fill(T* p, int64_t n, T v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) p[t] = v;
}

// include/devicearray.h:175-194: a(i,j,k) += b(i,j,k) over the shape
template <class T>
__global__ void k_add(View<T> a, View<T> b, Box bx, int64_t zc) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    a.at(i, j, k) += b.at(i, j, k);
}

// include/devicearray.h:211-237: max |a| over the shape
template <class T>
__global__ void k_norm(View<T> a, Box bx, unsigned long long* out) {
    int64_t j, k;
    double v = 0.0;
    if (column_point(bx, j, k))
        for (int64_t i = bx.x0; i < bx.x1; i++) v = fmax(v, fabs((double)a.at(i, j, k)));
    block_max_into(v, out);
}

// mixed precision: out(i,j,k) = (To)in(i,j,k) over the shape
template <class Ti, class To>
__global__ void k_cast(View<Ti> in, View<To> out, Box bx, int64_t zc) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    out.at(i, j, k) = (To)in.at(i, j, k);
}

// mixed precision: u(i,j,k) += (double)e(i,j,k) over the shape
__global__ void k_add_f32_to_f64(View<double> u, View<float> e, Box bx, int64_t zc) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    u.at(i, j, k) += (double)e.at(i, j, k);
}

// synthetic right-hand side: f(i,j,k) = sum_m A_m sin(a_m pi x) sin(b_m pi y)
// sin(c_m pi z) at x = i h, y = j h, z = k h, over the shape
struct Modes {
    int32_t n;
    int32_t abc[3 * MGP_MAX_MODES];
    double amp[MGP_MAX_MODES];
};

template <class T>
__global__ void k_fill_modes(View<T> f, Box bx, int64_t zc, Modes m, double h) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    const double x = (double)i * h, y = (double)j * h, z = (double)k * h;
    double s = 0.0;
    for (int q = 0; q < m.n; q++)
        s += m.amp[q] * sin(m.abc[3 * q] * M_PI * x) * sin(m.abc[3 * q + 1] * M_PI * y) * sin(m.abc[3 * q + 2] * M_PI * z);
    f.at(i, j, k) = (T)s;
}

// the shape of a halo-padded array into a dense C-order buffer
template <class T>
__global__ void k_pack(View<T> a, Box bx, int64_t zc, T* out) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    out[(i * a.n1 + j) * a.n2 + k] = a.at(i, j, k);
}

__global__ void k_zero_scalars(double* p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

// ----------------------------------------------------------- smoothers

// Colour of (i,j,k): reference = parity of the halo-inclusive linear index
// (include/gaussseidel.h:56,81; `lin0` is the linear index of (0,0,0));
// consistent = parity of i+j+k.
template <int MODE>
__device__ __forceinline__ int colour_of(int64_t lin0, int64_t s0, int64_t s1, int64_t i, int64_t j, int64_t k) {
    if (MODE == MGP_MODE_REFERENCE) return (int)((lin0 + i * s0 + j * s1 + k) & 1);
    return (int)((i + j + k) & 1);
}


// include/gaussseidel.h:28-94: one half-sweep in place, neighbours read from
// memory (the halo holds the ghosts written just before)
template <class T, int MODE>
__global__ void k_gs_half(View<T> u, View<T> f, Box bx, int64_t zc, T one_omega, T omega_sixth, T hsq, int colour,
                          int64_t lin0) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    if (colour_of<MODE>(lin0, u.s0, u.s1, i, j, k) != colour) return;
    const int64_t c = u.ix(i, j, k);
    T* p = u.p;
    const T fc = f.at(i, j, k);
    p[c] = relax_point<T, MODE>(p[c], p[c - u.s0], p[c + u.s0], p[c - u.s1], p[c + u.s1], p[c - 1], p[c + 1], fc,
                                one_omega, omega_sixth, hsq);
}

// nsweeps one-pass red-black sweeps of a level that fits in one workgroup's
// LDS, in one launch: the padded u (halos included) and f are staged in LDS;
// each sweep is ghost refresh, red half-sweep, ghost refresh, black
// half-sweep (include/relaxation.h:25-90), with barriers between, exactly
// what the per-sweep kernels compute; the box points are written to uo at
// the end.  Latency, not bandwidth, bounds the coarsest levels, and one
// launch replaces 2 nsweeps of them.  S is the arrays' storage type, T the
// arithmetic (and LDS) type: S = float with T = double is the FP64 solve of
// an FP32 coarsest level (values widen exactly on load, round once on store).
template <class S, class T, int MODE>
__global__ void __launch_bounds__(1024) k_rbgs_lds(View<S> ui, View<S> uo, View<S> f, Faces<S> F, Box bx,
                                                    int64_t nsweeps, T one_omega, T omega_sixth, T hsq, T two_h,
                                                    int64_t lin0) {
    constexpr int MAXP = 8;  // box points per thread: <= 8192 / 1024 for any level that fits
    extern __shared__ unsigned char lds_raw[];
    T* su = reinterpret_cast<T*>(lds_raw);
    // sizes below fit 32-bit indices (the level fits in 64 KiB)
    const int n0 = (int)ui.n0, n1 = (int)ui.n1, n2 = (int)ui.n2;
    // LDS layout: the shape with one ghost layer on every side, then f over
    // the shape.  Ghost cells are written by the refresh before every
    // half-sweep (Neumann faces) or never read (Dirichlet faces: the box
    // stops one short of them), so only the shape is staged.
    const int q2 = n2 + 2, q1 = n1 + 2;
    const int P = (n0 + 2) * q1 * q2;
    T* sf = su + P;
    auto L = [&](int i, int j, int k) { return ((i + 1) * q1 + (j + 1)) * q2 + (k + 1); };
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int t = tid; t < P; t += nt) su[t] = (T)0;
    __syncthreads();
    for (int t = tid; t < n0 * n1 * n2; t += nt) {
        const int i = t / (n1 * n2), j = (t / n2) % n1, k = t % n2;
        su[L(i, j, k)] = (T)ui.at(i, j, k);
        sf[t] = (T)f.at(i, j, k);
    }
    // this thread's box points: LDS index, f index, colour
    const int bn0 = (int)(bx.x1 - bx.x0), bn1 = (int)(bx.y1 - bx.y0), bn2 = (int)(bx.z1 - bx.z0);
    const int nbox = bn0 * bn1 * bn2;
    int pc[MAXP], pf[MAXP], pcol[MAXP];
#pragma unroll
    for (int m = 0; m < MAXP; m++) {
        const int t = tid + m * nt;
        pc[m] = -1;
        pf[m] = 0;
        pcol[m] = 2;
        if (t < nbox) {
            const int i = (int)bx.x0 + t / (bn1 * bn2), j = (int)bx.y0 + (t / bn2) % bn1, k = (int)bx.z0 + t % bn2;
            pc[m] = L(i, j, k);
            pf[m] = (i * n1 + j) * n2 + k;
            pcol[m] = colour_of<MODE>(lin0, ui.s0, ui.s1, i, j, k);
        }
    }
    const int s1l = q2, s0l = q1 * q2;
    // Neumann ghosts: sign*2h*g + mirror (include/neumann.h:112-130).  Each
    // thread owns at most one ghost cell per face (faces of a level this
    // kernel takes have at most blockDim cells): its LDS index, the mirror's
    // offset and the constant sign*2h*g are formed once, not per half-sweep.
    int gc[6], gd[6];
    T gv[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        gc[q] = -1;
        gd[q] = 0;
        gv[q] = (T)0;
        if (!F.neumann[q]) continue;
        const int a0 = (q == MGP_EAST || q == MGP_WEST) ? 1 : n0;
        const int a1 = (q == MGP_NORTH || q == MGP_SOUTH) ? 1 : n1;
        const int a2 = (q == MGP_TOP || q == MGP_BOTTOM) ? 1 : n2;
        const int t = tid;
        if (t >= a0 * a1 * a2) continue;
        const int i = t / (a1 * a2), j = (t / a2) % a1, k = t % a2;
        int gi = i, gj = j, gk = k, d = 0;
        T sign = (T)1.0;
        switch (q) {
        case MGP_EAST: gi = n0; d = -2 * s0l; break;
        case MGP_WEST: gi = -1; d = 2 * s0l; sign = (T)-1.0; break;
        case MGP_NORTH: gj = n1; d = -2 * s1l; break;
        case MGP_SOUTH: gj = -1; d = 2 * s1l; sign = (T)-1.0; break;
        case MGP_TOP: gk = n2; d = -2; break;
        default: gk = -1; d = 2; sign = (T)-1.0; break;
        }
        gc[q] = L(gi, gj, gk);
        gd[q] = d;
        gv[q] = sign * two_h * (T)F.g[q][t];
    }
    __syncthreads();
    for (int64_t s = 0; s < nsweeps; s++) {
        for (int colour = 0; colour < 2; colour++) {
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (gc[q] >= 0) su[gc[q]] = gv[q] + su[gc[q] + gd[q]];
            __syncthreads();
#pragma unroll
            for (int m = 0; m < MAXP; m++) {
                if (pcol[m] != colour) continue;
                const int c = pc[m];
                su[c] = relax_point<T, MODE>(su[c], su[c - s0l], su[c + s0l], su[c - s1l], su[c + s1l], su[c - 1],
                                             su[c + 1], sf[pf[m]], one_omega, omega_sixth, hsq);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int m = 0; m < MAXP; m++) {
        if (pc[m] < 0) continue;
        const int t = tid + m * nt;
        const int i = (int)bx.x0 + t / (bn1 * bn2), j = (int)bx.y0 + (t / bn2) % bn1, k = (int)bx.z0 + t % bn2;
        uo.at(i, j, k) = (S)su[pc[m]];
    }
    // out of place: the shape points outside the box as loaded, so u_out is a
    // complete image of the in-place result (the boundary points' own history,
    // signs of zero included, follows whichever buffer ends up current)
    if (uo.p != ui.p) {
        for (int t = tid; t < n0 * n1 * n2; t += nt) {
            const int i = t / (n1 * n2), j = (t / n2) % n1, k = t % n2;
            if (i >= bx.x0 && i < bx.x1 && j >= bx.y0 && j < bx.y1 && k >= bx.z0 && k < bx.z1) continue;
            uo.at(i, j, k) = (S)su[L(i, j, k)];
        }
    }
}

// ------------------------------------------------- coarsest-level CG solve
// The coarsest level of a cycle solved by conjugate gradients in FP64 in one
// workgroup (mgp_cg_solve_*; no counterpart in the reference, whose coarsest
// level gets the same few smoothing passes as the others, include/vcycle.h:
// 27-31).  Smoothing passes leave the smooth error of a coarsest level with a
// long axis and few Dirichlet faces nearly untouched (a 33-point axis with one
// Dirichlet end needs thousands of Gauss-Seidel passes), which stalls the
// cycle; CG reduces it in O(sqrt(condition)) steps.
//
// With res = f - L u (Neumann ghosts sign*2h*g + mirror, Dirichlet boundary
// nodes fixed), the correction d solves L0 d = res, L0 the operator with
// homogeneous data.  L0 is not symmetric at Neumann boundary nodes (the ghost
// doubles the inward coupling), but W L0 is, with W = 1/2 per Neumann
// boundary coordinate of a point; CG runs on the SPD system -W L0 d = -W res
// from d = 0 until |r| <= rtol |r0| or max_iters, then u += d.  The search
// direction p lives in LDS over the shape padded by one layer (ghost layer:
// the mirror value, Dirichlet nodes: 0); each thread keeps the d, r and
// weights of its points in registers.  Reductions are fixed-order trees, so
// the result does not depend on timing.
__device__ __forceinline__ double block_sum(double v, double* part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) part[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nwaves; w++) t += part[w];
    __syncthreads();  // part is reused by the next reduction
    return t;
}

template <class S>
__global__ void __launch_bounds__(1024) k_cg_lds(View<S> u, View<S> f, Faces<S> F, Box bx, int64_t max_iters,
                                                  double rtol, double two_h, double inv_h2) {
    constexpr int MAXP = 8;  // box points per thread (<= 8192 / 1024)
    extern __shared__ unsigned char lds_raw[];
    __shared__ double part[16];
    double* sp = reinterpret_cast<double*>(lds_raw);
    const int n0 = (int)u.n0, n1 = (int)u.n1, n2 = (int)u.n2;
    const int q2 = n2 + 2, q1 = n1 + 2;
    const int P = (n0 + 2) * q1 * q2;
    auto L = [&](int i, int j, int k) { return ((i + 1) * q1 + (j + 1)) * q2 + (k + 1); };
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int t = tid; t < P; t += nt) sp[t] = 0.0;
    const int bn0 = (int)(bx.x1 - bx.x0), bn1 = (int)(bx.y1 - bx.y0), bn2 = (int)(bx.z1 - bx.z0);
    const int nbox = bn0 * bn1 * bn2;
    int pc[MAXP];
    double w[MAXP], d[MAXP], r[MAXP];
    double rr = 0.0;
#pragma unroll
    for (int m = 0; m < MAXP; m++) {
        const int t = tid + m * nt;
        pc[m] = -1;
        w[m] = d[m] = r[m] = 0.0;
        if (t >= nbox) continue;
        const int i = (int)bx.x0 + t / (bn1 * bn2), j = (int)bx.y0 + (t / bn2) % bn1, k = (int)bx.z0 + t % bn2;
        pc[m] = L(i, j, k);
        // weight: 1/2 per coordinate on a Neumann face
        double wt = 1.0;
        if ((i == 0 && F.neumann[MGP_WEST]) || (i == n0 - 1 && F.neumann[MGP_EAST])) wt *= 0.5;
        if ((j == 0 && F.neumann[MGP_SOUTH]) || (j == n1 - 1 && F.neumann[MGP_NORTH])) wt *= 0.5;
        if ((k == 0 && F.neumann[MGP_BOTTOM]) || (k == n2 - 1 && F.neumann[MGP_TOP])) wt *= 0.5;
        w[m] = wt;
        // res = f - L u with the stored boundary values and Neumann ghosts
        const double uc = (double)u.at(i, j, k);
        double sum = 0.0;
#pragma unroll
        for (int ax = 0; ax < 3; ax++)
#pragma unroll
            for (int dir = -1; dir <= 1; dir += 2) {
                int a[3] = {i, j, k};
                const int n[3] = {n0, n1, n2};
                a[ax] += dir;
                double nb;
                if (a[ax] >= 0 && a[ax] < n[ax]) {
                    nb = (double)u.at(a[0], a[1], a[2]);
                } else {
                    const int face = ax == 0 ? (dir > 0 ? MGP_EAST : MGP_WEST)
                                             : ax == 1 ? (dir > 0 ? MGP_NORTH : MGP_SOUTH) : (dir > 0 ? MGP_TOP : MGP_BOTTOM);
                    a[ax] -= 2 * dir;
                    const int gi = ax == 0 ? j * n2 + k : ax == 1 ? i * n2 + k : i * n1 + j;
                    const double g = F.g[face] ? (double)F.g[face][gi] : 0.0;
                    nb = (double)dir * two_h * g + (double)u.at(a[0], a[1], a[2]);
                }
                sum += nb - uc;
            }
        const double res = fma(-inv_h2, sum, (double)f.at(i, j, k));
        r[m] = -wt * res;  // c = -W res; r0 = c with d = 0
        rr += r[m] * r[m];
    }
    __syncthreads();  // sp zeroed before p is written
#pragma unroll
    for (int m = 0; m < MAXP; m++)
        if (pc[m] >= 0) sp[pc[m]] = r[m];
    rr = block_sum(rr, part);
    const double stop = rtol * rtol * rr;
    // ghost cells of the padded shape on Neumann faces: p(ghost) = p(mirror);
    // each thread owns up to GMAX cells per face (faces of <= GMAX * 1024
    // cells; a level that fits the LDS has faces of at most ~2700)
    constexpr int GMAX = 4;
    const int s0l = q1 * q2, s1l = q2;
    int gc[6][GMAX], gd[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        gd[q] = q == MGP_EAST ? -2 * s0l : q == MGP_WEST ? 2 * s0l : q == MGP_NORTH ? -2 * s1l
              : q == MGP_SOUTH ? 2 * s1l : q == MGP_TOP ? -2 : 2;
        const int a0 = (q == MGP_EAST || q == MGP_WEST) ? 1 : n0;
        const int a1 = (q == MGP_NORTH || q == MGP_SOUTH) ? 1 : n1;
        const int a2 = (q == MGP_TOP || q == MGP_BOTTOM) ? 1 : n2;
#pragma unroll
        for (int g = 0; g < GMAX; g++) {
            const int t = tid + g * nt;
            gc[q][g] = -1;
            if (!F.neumann[q] || t >= a0 * a1 * a2) continue;
            const int i = t / (a1 * a2), j = (t / a2) % a1, k = t % a2;
            switch (q) {
            case MGP_EAST: gc[q][g] = L(n0, j, k); break;
            case MGP_WEST: gc[q][g] = L(-1, j, k); break;
            case MGP_NORTH: gc[q][g] = L(i, n1, k); break;
            case MGP_SOUTH: gc[q][g] = L(i, -1, k); break;
            case MGP_TOP: gc[q][g] = L(i, j, n2); break;
            default: gc[q][g] = L(i, j, -1); break;
            }
        }
    }
    for (int64_t it = 0; it < max_iters && rr > stop && rr > 0.0; it++) {
#pragma unroll
        for (int q = 0; q < 6; q++)
#pragma unroll
            for (int g = 0; g < GMAX; g++)
                if (gc[q][g] >= 0) sp[gc[q][g]] = sp[gc[q][g] + gd[q]];
        __syncthreads();
        double qv[MAXP], pq = 0.0;
#pragma unroll
        for (int m = 0; m < MAXP; m++) {
            qv[m] = 0.0;
            if (pc[m] < 0) continue;
            const int c = pc[m];
            const double pcv = sp[c];
            const double sum = ((sp[c - s0l] - pcv) + (sp[c + s0l] - pcv)) + ((sp[c - s1l] - pcv) + (sp[c + s1l] - pcv)) +
                               ((sp[c - 1] - pcv) + (sp[c + 1] - pcv));
            qv[m] = -w[m] * inv_h2 * sum;  // (-W L0 p)
            pq += pcv * qv[m];
        }
        pq = block_sum(pq, part);
        if (!(pq > 0.0)) break;
        const double alpha = rr / pq;
        double rn = 0.0;
#pragma unroll
        for (int m = 0; m < MAXP; m++) {
            if (pc[m] < 0) continue;
            d[m] += alpha * sp[pc[m]];
            r[m] -= alpha * qv[m];
            rn += r[m] * r[m];
        }
        rn = block_sum(rn, part);
        const double beta = rn / rr;
        rr = rn;
#pragma unroll
        for (int m = 0; m < MAXP; m++)
            if (pc[m] >= 0) sp[pc[m]] = r[m] + beta * sp[pc[m]];
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < MAXP; m++) {
        if (pc[m] < 0) continue;
        const int t = tid + m * nt;
        const int i = (int)bx.x0 + t / (bn1 * bn2), j = (int)bx.y0 + (t / bn2) % bn1, k = (int)bx.z0 + t % bn2;
        u.at(i, j, k) = (S)((double)u.at(i, j, k) + d[m]);
    }
}

// include/jacobi.h:39-82: u = J(v), out of place; the reference writes
// u = (1-w)v and then adds, which is the same two roundings
template <class T, int MODE>
__global__ void k_jacobi(View<T> u, View<T> v, View<T> f, Box bx, int64_t zc, T one_omega, T omega_sixth, T hsq) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    const int64_t c = v.ix(i, j, k);
    const T* p = v.p;
    u.at(i, j, k) = relax_point<T, MODE>(p[c], p[c - v.s0], p[c + v.s0], p[c - v.s1], p[c + v.s1], p[c - 1], p[c + 1],
                                         f.at(i, j, k), one_omega, omega_sixth, hsq);
}

// ------------------------------------------------------------- residual


template <class T, int MODE>
__global__ void k_residual(View<T> u, View<T> f, View<T> r, Box bx, int64_t zc, T hsq, T inv_h2) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    r.at(i, j, k) = residual_point<T, MODE>(u, u.ix(i, j, k), f.at(i, j, k), hsq, inv_h2);
}

// Residual on the box plus max |r| over r's whole shape in the same pass
// (the columns cover the shape; points outside the box only contribute).
template <class T, int MODE>
__global__ void k_residual_norm(View<T> u, View<T> f, View<T> r, Box shape, Box bx, T hsq, T inv_h2,
                                unsigned long long* out) {
    int64_t j, k;
    double v = 0.0;
    if (column_point(shape, j, k)) {
        const bool col_in = j >= bx.y0 && j < bx.y1 && k >= bx.z0 && k < bx.z1;
        for (int64_t i = shape.x0; i < shape.x1; i++) {
            T rv;
            if (col_in && i >= bx.x0 && i < bx.x1) {
                rv = residual_point<T, MODE>(u, u.ix(i, j, k), f.at(i, j, k), hsq, inv_h2);
                r.at(i, j, k) = rv;
            } else {
                rv = r.at(i, j, k);
            }
            v = fmax(v, fabs((double)rv));
        }
    }
    block_max_into(v, out);
}

// ------------------------------------------------------- grid transfers

// include/injection.h:47-66
template <class T>
__global__ void k_inject(View<T> in, View<T> out, Box bx, int64_t zc) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    out.at(i, j, k) = in.at(2 * i, 2 * j, 2 * k);
}

// include/fullweighting.h:59-121
template <class T>
__global__ void k_fw(View<T> in, View<T> out, Box bx, int64_t zc) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    out.at(i, j, k) = fw_point<T>(in, i, j, k);
}


template <class T, bool ADD>
__global__ void k_prolong(View<T> coarse, View<T> fine, Box bx, int64_t zc) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    const T v = trilinear<T>(coarse, i, j, k);
    if (ADD)
        fine.at(i, j, k) += v;  // include/devicearray.h:190
    else
        fine.at(i, j, k) = v;
}

// --------------------------------------------------- boundary conditions

// Offsets of face `loc` in the domain: the face array's (i,j,k) maps to the
// boundary node (i+o0, j+o1, k+o2) (include/dirichlet.h:53-68)
__device__ __forceinline__ void face_geometry(int loc, int64_t n0, int64_t n1, int64_t n2, int64_t& f0, int64_t& f1,
                                              int64_t& f2) {
    f0 = (loc == MGP_EAST || loc == MGP_WEST) ? 1 : n0;
    f1 = (loc == MGP_NORTH || loc == MGP_SOUTH) ? 1 : n1;
    f2 = (loc == MGP_TOP || loc == MGP_BOTTOM) ? 1 : n2;
}

template <class T>
__global__ void k_bc_dirichlet(View<T> u, int loc, const T* g) {
    int64_t f0, f1, f2;
    face_geometry(loc, u.n0, u.n1, u.n2, f0, f1, f2);
    const int64_t n = f0 * f1 * f2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = t / (f1 * f2), j = (t / f2) % f1, k = t % f2;
    const int64_t o0 = loc == MGP_EAST ? u.n0 - 1 : 0;
    const int64_t o1 = loc == MGP_NORTH ? u.n1 - 1 : 0;
    const int64_t o2 = loc == MGP_TOP ? u.n2 - 1 : 0;
    u.at(i + o0, j + o1, k + o2) = g[t];
}

// include/neumann.h:63-131
template <class T>
__global__ void k_bc_neumann(View<T> u, int loc, const T* g, T two_h) {
    int64_t f0, f1, f2;
    face_geometry(loc, u.n0, u.n1, u.n2, f0, f1, f2);
    const int64_t n = f0 * f1 * f2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = t / (f1 * f2), j = (t / f2) % f1, k = t % f2;
    int64_t a = i, b = j, c = k, d0 = 0, d1 = 0, d2 = 0;
    T sign = 1.0;
    switch (loc) {
    case MGP_EAST: a = u.n0; d0 = -2; break;
    case MGP_WEST: a = -1; d0 = 2; sign = -1.0; break;
    case MGP_NORTH: b = u.n1; d1 = -2; break;
    case MGP_SOUTH: b = -1; d1 = 2; sign = -1.0; break;
    case MGP_TOP: c = u.n2; d2 = -2; break;
    default: c = -1; d2 = 2; sign = -1.0; break;
    }
    const T gv = g ? g[t] : (T)0.0;
    u.at(a, b, c) = sign * two_h * gv + u.at(a + d0, b + d1, c + d2);
}

// The ghosts of every Neumann face in one launch: thread t belongs to the
// face whose range [start[q], start[q+1]) holds it and does what k_bc_neumann
// does for that face.  Each face writes only its own ghost layer (the face's
// points, edges of the halo excluded) and reads points inside the shape, so
// the faces do not depend on one another and the result equals the per-face
// refreshes in any order.
template <class T>
struct NeumannFaces {
    const T* g[6];
    int64_t start[7];  // prefix of the Neumann faces' point counts (0 for the others)
    T two_h;
};

template <class T>
__global__ void k_bc_neumann_faces(View<T> u, NeumannFaces<T> F) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= F.start[6]) return;
    int loc = 0;
    while (t >= F.start[loc + 1]) loc++;
    int64_t f0, f1, f2;
    face_geometry(loc, u.n0, u.n1, u.n2, f0, f1, f2);
    const int64_t p = t - F.start[loc];
    const int64_t i = p / (f1 * f2), j = (p / f2) % f1, k = p % f2;
    int64_t a = i, b = j, c = k, d0 = 0, d1 = 0, d2 = 0;
    T sign = 1.0;
    switch (loc) {
    case MGP_EAST: a = u.n0; d0 = -2; break;
    case MGP_WEST: a = -1; d0 = 2; sign = -1.0; break;
    case MGP_NORTH: b = u.n1; d1 = -2; break;
    case MGP_SOUTH: b = -1; d1 = 2; sign = -1.0; break;
    case MGP_TOP: c = u.n2; d2 = -2; break;
    default: c = -1; d2 = 2; sign = -1.0; break;
    }
    const T gv = F.g[loc] ? F.g[loc][p] : (T)0.0;
    u.at(a, b, c) = sign * F.two_h * gv + u.at(a + d0, b + d1, c + d2);
}

// include/neumann.h:146-252: coarse face data = R(fine face data), then minus
// the one-sided difference of the fine solution at the face divided by h_fine.
template <class T>
__global__ void k_bc_restrict_neumann(int restriction, View<T> uf, int loc, View<T> gf, View<T> gc, T c1, T c2, T c3,
                                      T hf) {
    const int64_t n = gc.n0 * gc.n1 * gc.n2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = t / (gc.n1 * gc.n2), j = (t / gc.n2) % gc.n1, k = t % gc.n2;
    T v;
    if (restriction == MGP_FULL_WEIGHTING) {
        const int64_t n0 = gf.n0, n1 = gf.n1, n2 = gf.n2;
        const int64_t ii = min(n0 - 1, 2 * i), jj = min(n1 - 1, 2 * j), kk = min(n2 - 1, 2 * k);
        const int64_t im = max((int64_t)0, ii - 1), ip = min(n0 - 1, ii + 1);
        const int64_t jm = max((int64_t)0, jj - 1), jp = min(n1 - 1, jj + 1);
        const int64_t km = max((int64_t)0, kk - 1), kp = min(n2 - 1, kk + 1);
#define A(a, b, c) gf.at(a, b, c)
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
        v = tmp / 64.0;
    } else {
        v = gf.at(2 * i, 2 * j, 2 * k);
    }
    int64_t o0 = 0, o1 = 0, o2 = 0, d0 = 0, d1 = 0, d2 = 0;
    switch (loc) {
    case MGP_EAST: o0 = uf.n0 / 2; d0 = -1; break;
    case MGP_WEST: d0 = 1; break;
    case MGP_NORTH: o1 = uf.n1 / 2; d1 = -1; break;
    case MGP_SOUTH: d1 = 1; break;
    case MGP_TOP: o2 = uf.n2 / 2; d2 = -1; break;
    default: d2 = 1; break;
    }
    const int64_t a = 2 * (i + o0), b = 2 * (j + o1), c = 2 * (k + o2);
    v -= (c1 * uf.at(a, b, c) + c2 * uf.at(a + d0, b + d1, c + d2) + c3 * uf.at(a + 2 * d0, b + 2 * d1, c + 2 * d2)) /
         hf;
    gc.at(i, j, k) = v;
}

// ======================================================= host-side wrappers

// ============================================== FP16 storage, FP32 arithmetic
//
// The coarse levels of an FP32 cycle may be stored as IEEE binary16 (an
// addition; the reference computes in one type T throughout).  Values are
// loaded as float, combined with the consistent-mode point arithmetic of the
// FP32 kernels, and rounded to binary16 (nearest even) when stored.  All
// faces Dirichlet: the boxes stay one point inside the shape.
//
// Scaling: a level's right-hand side is the restricted residual of the level
// above times sigma = 2^-e, where max |r| = m 2^e with m in [0.5, 1) (read
// on the device from the norm the residual kernel wrote), so restricted
// values stay near 1 whatever the refinement has reduced the residual to;
// the prolongation multiplies the correction by 2^e again.  Powers of two
// scale exactly, so the scaled problem is the same linear problem.

// v rounded to float, then to binary16 (nearest even): the empty asm keeps
// the compiler from folding the operation that produced v into the
// conversion (v_fma_mixlo_f16 and the like round once, straight to binary16)
__device__ __forceinline__ _Float16 to_half(float v) {
    asm volatile("" : "+v"(v));
    return (_Float16)v;
}

struct HalfView {
    _Float16* p;  // element (0,0,0) of the shape
    int64_t n0, n1, n2, s0, s1;
    explicit HalfView(const View<_Float16>& v) : p(v.p), n0(v.n0), n1(v.n1), n2(v.n2), s0(v.s0), s1(v.s1) {}
    __device__ __forceinline__ int64_t ix(int64_t i, int64_t j, int64_t k) const { return i * s0 + j * s1 + k; }
    __device__ __forceinline__ float ld(int64_t c) const { return (float)p[c]; }
    __device__ __forceinline__ float at(int64_t i, int64_t j, int64_t k) const { return (float)p[ix(i, j, k)]; }
    __device__ __forceinline__ void st(int64_t c, float v) const { p[c] = to_half(v); }
};

__device__ __forceinline__ void add_into(const View<float>& a, int64_t i, int64_t j, int64_t k, float v) {
    a.at(i, j, k) += v;
}
__device__ __forceinline__ void add_into(const HalfView& a, int64_t i, int64_t j, int64_t k, float v) {
    const int64_t c = a.ix(i, j, k);
    a.st(c, a.ld(c) + v);
}

// 2^-e (or 2^e with `inverse`) for the max |r| in *norm; 1 for a zero or
// non-finite norm
__device__ __forceinline__ float pow2_scale(const double* norm, bool inverse) {
    const double n = *norm;
    if (!(n > 0.0) || !(n < 1e300)) return 1.0f;
    int e;
    (void)frexp(n, &e);
    e = min(120, max(-120, e));
    return ldexpf(1.0f, inverse ? e : -e);
}

// consistent-mode residual_point on FP16 storage
__device__ __forceinline__ float residual_point_h(const HalfView& u, int64_t c, float fc, float inv_h2) {
    const float uc = u.ld(c);
    const float s = ((u.ld(c - u.s0) - uc) + (u.ld(c + u.s0) - uc)) + ((u.ld(c - u.s1) - uc) + (u.ld(c + u.s1) - uc)) +
                    ((u.ld(c - 1) - uc) + (u.ld(c + 1) - uc));
    return fmaf(-inv_h2, s, fc);
}

__global__ void k_fill_h(_Float16* p, int64_t n, float v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const _Float16 hv = to_half(v);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) p[t] = hv;
}

// one red (colour 0) or black half-sweep in place, colour = parity of i+j+k
__global__ void k_gs_half_h(HalfView u, HalfView f, Box bx, int64_t zc, float omega_sixth, float hsq, int colour) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    if ((int)((i + j + k) & 1) != colour) return;
    const int64_t c = u.ix(i, j, k);
    u.st(c, relax_point<float, MGP_MODE_CONSISTENT>(u.ld(c), u.ld(c - u.s0), u.ld(c + u.s0), u.ld(c - u.s1),
                                                    u.ld(c + u.s1), u.ld(c - 1), u.ld(c + 1), f.at(i, j, k), 0.0f,
                                                    omega_sixth, hsq));
}

// nsweeps red-black sweeps of a level that fits one workgroup's LDS, in one
// launch: u and f staged as float, every update rounded to binary16 as the
// per-colour kernel stores it, so the result equals 2 nsweeps launches of
// k_gs_half_h
__global__ void __launch_bounds__(1024) k_sweeps_lds_h(HalfView u, HalfView f, Box bx, int64_t nsweeps,
                                                        float omega_sixth, float hsq) {
    constexpr int MAXP = 8;  // box points per thread (<= 8192 points)
    extern __shared__ unsigned char lds_raw[];
    float* su = reinterpret_cast<float*>(lds_raw);
    const int n0 = (int)u.n0, n1 = (int)u.n1, n2 = (int)u.n2;
    const int N = n0 * n1 * n2;
    float* sf = su + N;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int t = tid; t < N; t += nt) {
        const int i = t / (n1 * n2), j = (t / n2) % n1, k = t % n2;
        su[t] = u.at(i, j, k);
        sf[t] = f.at(i, j, k);
    }
    const int bn0 = (int)(bx.x1 - bx.x0), bn1 = (int)(bx.y1 - bx.y0), bn2 = (int)(bx.z1 - bx.z0);
    const int nbox = bn0 * bn1 * bn2;
    int pc[MAXP], pcol[MAXP];
#pragma unroll
    for (int m = 0; m < MAXP; m++) {
        const int t = tid + m * nt;
        pc[m] = -1;
        pcol[m] = 2;
        if (t < nbox) {
            const int i = (int)bx.x0 + t / (bn1 * bn2), j = (int)bx.y0 + (t / bn2) % bn1, k = (int)bx.z0 + t % bn2;
            pc[m] = (i * n1 + j) * n2 + k;
            pcol[m] = (i + j + k) & 1;
        }
    }
    const int s1l = n2, s0l = n1 * n2;
    __syncthreads();
    for (int64_t s = 0; s < nsweeps; s++) {
        for (int colour = 0; colour < 2; colour++) {
#pragma unroll
            for (int m = 0; m < MAXP; m++) {
                if (pcol[m] != colour) continue;
                const int c = pc[m];
                const float v = relax_point<float, MGP_MODE_CONSISTENT>(su[c], su[c - s0l], su[c + s0l], su[c - s1l],
                                                                        su[c + s1l], su[c - 1], su[c + 1], sf[c],
                                                                        0.0f, omega_sixth, hsq);
                su[c] = (float)to_half(v);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int m = 0; m < MAXP; m++) {
        if (pc[m] < 0) continue;
        const int c = pc[m];
        const int i = c / (n1 * n2), j = (c / n2) % n1, k = c % n2;
        u.st(u.ix(i, j, k), su[c]);
    }
}

// r = f - L u on the box (stored as binary16) and max |r| of the stored
// values over r's whole shape into *out (columns over the shape, as
// k_residual_norm)
__global__ void k_residual_norm_h(HalfView u, HalfView f, HalfView r, Box shape, Box bx, float inv_h2,
                                  unsigned long long* out) {
    int64_t j, k;
    double v = 0.0;
    if (column_point(shape, j, k)) {
        const bool col_in = j >= bx.y0 && j < bx.y1 && k >= bx.z0 && k < bx.z1;
        for (int64_t i = shape.x0; i < shape.x1; i++) {
            const int64_t c = r.ix(i, j, k);
            if (col_in && i >= bx.x0 && i < bx.x1) r.st(c, residual_point_h(u, u.ix(i, j, k), f.at(i, j, k), inv_h2));
            v = fmax(v, fabs((double)r.ld(c)));
        }
    }
    if (out) block_max_into(v, out);
}

// coarse f = sigma * FW(fine r), fine r in FP32 (the last FP32 level) or FP16
template <class V>
__global__ void k_fw_scaled_h(V in, HalfView out, Box bx, int64_t zc, const double* norm) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    const float sigma = pow2_scale(norm, false);
    out.st(out.ix(i, j, k), fw_point<float, V>(in, i, j, k) * sigma);
}

// fine u += P(coarse u) / sigma over the fine shape, fine u in FP32 or FP16
template <class V>
__global__ void k_prolong_add_scaled_h(HalfView coarse, V fine, Box bx, int64_t zc, const double* norm) {
    int64_t i, j, k;
    if (!box_point(bx, zc, i, j, k)) return;
    const float inv = pow2_scale(norm, true);
    add_into(fine, i, j, k, trilinear<float, HalfView>(coarse, i, j, k) * inv);
}


static inline int launch_status() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? MGP_OK : (int)e;
}

static inline bool valid_array(const mgp_array* a) {
    if (!a || !a->data) return false;
    for (int d = 0; d < 3; d++)
        if (a->shape[d] < 1) return false;
    for (int q = 0; q < 6; q++)
        if (a->halo[q] < 0) return false;
    return a->stride[2] == 1 && a->stride[1] >= a->shape[2] + a->halo[MGP_TOP] + a->halo[MGP_BOTTOM] &&
           a->stride[0] >= a->stride[1] * (a->shape[1] + a->halo[MGP_NORTH] + a->halo[MGP_SOUTH]);
}

static inline bool same_shape(const mgp_array* a, const mgp_array* b) {
    return a->shape[0] == b->shape[0] && a->shape[1] == b->shape[1] && a->shape[2] == b->shape[2];
}

static inline bool valid_box(const mgp_array* a, const int64_t b[6]) {
    // the stencil reads one point beyond the box on each side: inside the
    // shape, or in a halo layer that exists
    const int64_t lo[3] = {a->halo[MGP_WEST], a->halo[MGP_SOUTH], a->halo[MGP_BOTTOM]};
    const int64_t hi[3] = {a->halo[MGP_EAST], a->halo[MGP_NORTH], a->halo[MGP_TOP]};
    for (int d = 0; d < 3; d++) {
        if (b[2 * d] >= b[2 * d + 1]) continue;
        if (b[2 * d] - 1 < -lo[d] || b[2 * d + 1] > a->shape[d] + hi[d]) return false;
    }
    return true;
}

static inline int64_t lin0_of(const mgp_array* a) {
    return a->halo[MGP_WEST] * a->stride[0] + a->halo[MGP_SOUTH] * a->stride[1] + a->halo[MGP_BOTTOM];
}

static inline Box shape_box(const mgp_array* a) { return Box{0, a->shape[0], 0, a->shape[1], 0, a->shape[2]}; }

template <class T>
static int fill(const mgp_array* a, double value, void* stream) {
    if (!a || !a->data) return MGP_ERR_ARG;
    const int64_t n = a->stride[0] * (a->shape[0] + a->halo[MGP_EAST] + a->halo[MGP_WEST]);
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (T*)a->data, n, (T)value);
    return launch_status();
}

template <class T>
static int add(const mgp_array* a, const mgp_array* b, void* stream) {
    if (!valid_array(a) || !valid_array(b) || !same_shape(a, b)) return MGP_ERR_ARG;
    const Box bx = shape_box(a);
    const BoxGrid g = box_grid(bx);
    hipLaunchKernelGGL(k_add<T>, g.grid, g.block, 0, (hipStream_t)stream, view_of<T>(*a), view_of<T>(*b), bx,
                       g.zchunks);
    return launch_status();
}

template <class T>
static int norm_inf(const mgp_array* a, double* out, void* stream) {
    if (!valid_array(a) || !out) return MGP_ERR_ARG;
    const Box bx = shape_box(a);
    const BoxGrid g = column_grid(bx);
    hipLaunchKernelGGL(k_norm<T>, g.grid, g.block, 0, (hipStream_t)stream, view_of<T>(*a), bx,
                       (unsigned long long*)out);
    return launch_status();
}

// constants exactly as the reference forms them from T omega and double h
// (include/gaussseidel.h:50-52, jacobi.h:42-44)

template <class T>
static int rbgs_half(int mode, const mgp_array* u, const mgp_array* f, const int64_t b[6], double omega, double h,
                     int colour, void* stream) {
    if (!valid_array(u) || !valid_array(f) || !same_shape(u, f) || !valid_box(u, b) || (colour & ~1)) return MGP_ERR_ARG;
    const Box bx = box_of(b);
    if (box_empty(bx)) return MGP_OK;
    const BoxGrid g = box_grid(bx);
    const RelaxConsts<T> c(omega, h);
    if (mode == MGP_MODE_REFERENCE)
        hipLaunchKernelGGL((k_gs_half<T, MGP_MODE_REFERENCE>), g.grid, g.block, 0, (hipStream_t)stream, view_of<T>(*u),
                           view_of<T>(*f), bx, g.zchunks, c.one_omega, c.omega_sixth, c.hsq, colour, lin0_of(u));
    else if (mode == MGP_MODE_CONSISTENT)
        hipLaunchKernelGGL((k_gs_half<T, MGP_MODE_CONSISTENT>), g.grid, g.block, 0, (hipStream_t)stream,
                           view_of<T>(*u), view_of<T>(*f), bx, g.zchunks, c.one_omega, c.omega_sixth, c.hsq, colour,
                           lin0_of(u));
    else
        return MGP_ERR_ARG;
    return launch_status();
}

static inline bool checkerboard(int mode, const mgp_array* u) {
    return mode == MGP_MODE_CONSISTENT || ((u->stride[0] & 1) && (u->stride[1] & 1));
}

template <class T>
static int rbgs_sweep(int mode, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* faces,
                      const int64_t b[6], double omega, void* stream) {
    if (!valid_array(ui) || !valid_array(uo) || !valid_array(f) || !faces || !same_shape(ui, uo) ||
        !same_shape(ui, f) || (mode != MGP_MODE_REFERENCE && mode != MGP_MODE_CONSISTENT))
        return MGP_ERR_ARG;
    if (!checkerboard(mode, ui) || lin0_of(ui) != lin0_of(uo) || ui->stride[0] != uo->stride[0] ||
        ui->stride[1] != uo->stride[1])
        return MGP_ERR_ARG;
    const Box bx = box_of(b);
    for (int d = 0; d < 3; d++)
        if (b[2 * d] < 0 || b[2 * d + 1] > ui->shape[d]) return MGP_ERR_ARG;
    // a box point on a face with a Dirichlet type would need a ghost beyond it
    if ((bx.x0 == 0 && faces->type[MGP_WEST] != MGP_NEUMANN) ||
        (bx.x1 == ui->shape[0] && faces->type[MGP_EAST] != MGP_NEUMANN) ||
        (bx.y0 == 0 && faces->type[MGP_SOUTH] != MGP_NEUMANN) ||
        (bx.y1 == ui->shape[1] && faces->type[MGP_NORTH] != MGP_NEUMANN) ||
        (bx.z0 == 0 && faces->type[MGP_BOTTOM] != MGP_NEUMANN) ||
        (bx.z1 == ui->shape[2] && faces->type[MGP_TOP] != MGP_NEUMANN))
        return MGP_ERR_ARG;
    // an empty box out of place would leave u_out without the copy of u_in
    if (box_empty(bx)) return ui->data == uo->data ? MGP_OK : MGP_ERR_ARG;
    return march_sweep<T>(mode, ui, uo, f, faces, b, omega, march_xchunk(b), stream);
}

// two sweeps in one pass, u_out = GS(GS(u_in)), u_out a different buffer;
// with coarse, u_out = GS(GS(u_in + P(coarse))) (consistent mode)
template <class T>
static int rbgs_sweep2(int mode, const mgp_array* ui, const mgp_array* uo, const mgp_array* f,
                       const mgp_faces* faces, const int64_t b[6], double omega, void* stream,
                       const mgp_array* coarse = nullptr) {
    if (!valid_array(ui) || !valid_array(uo) || !valid_array(f) || !faces || !same_shape(ui, uo) ||
        !same_shape(ui, f) || ui->data == uo->data || (mode != MGP_MODE_REFERENCE && mode != MGP_MODE_CONSISTENT))
        return MGP_ERR_ARG;
    if (coarse) {
        if (mode != MGP_MODE_CONSISTENT || !valid_array(coarse) || coarse->data == ui->data ||
            coarse->data == uo->data)
            return MGP_ERR_ARG;
        for (int d = 0; d < 3; d++)  // exact 2:1 nesting (trilinearinterpolation.h:31-37)
            if (2 * (coarse->shape[d] - 1) + 1 != ui->shape[d]) return MGP_ERR_ARG;
    }
    if (!checkerboard(mode, ui) || lin0_of(ui) != lin0_of(uo) || ui->stride[0] != uo->stride[0] ||
        ui->stride[1] != uo->stride[1])
        return MGP_ERR_ARG;
    const Box bx = box_of(b);
    for (int d = 0; d < 3; d++)
        if (b[2 * d] < 0 || b[2 * d + 1] > ui->shape[d]) return MGP_ERR_ARG;
    if ((bx.x0 == 0 && faces->type[MGP_WEST] != MGP_NEUMANN) ||
        (bx.x1 == ui->shape[0] && faces->type[MGP_EAST] != MGP_NEUMANN) ||
        (bx.y0 == 0 && faces->type[MGP_SOUTH] != MGP_NEUMANN) ||
        (bx.y1 == ui->shape[1] && faces->type[MGP_NORTH] != MGP_NEUMANN) ||
        (bx.z0 == 0 && faces->type[MGP_BOTTOM] != MGP_NEUMANN) ||
        (bx.z1 == ui->shape[2] && faces->type[MGP_TOP] != MGP_NEUMANN))
        return MGP_ERR_ARG;
    for (int q = 0; q < 6; q++)
        if (faces->type[q] == MGP_NEUMANN && !faces->g[q]) return MGP_ERR_ARG;
    if (box_empty(bx)) return MGP_ERR_ARG;  // u_out would not receive the copy of u_in
    return march_sweep2<T>(mode, ui, uo, f, faces, b, omega, stream, coarse);
}

// two sweeps of u = 0 into u_out (consistent mode): the checks of rbgs_sweep2
// with u_out as the layout of the (unread) input
template <class T>
static int rbgs_sweep2_from_zero(int mode, const mgp_array* uo, const mgp_array* f, const mgp_faces* faces,
                                 const int64_t b[6], double omega, void* stream) {
    if (!valid_array(uo) || !valid_array(f) || !faces || !same_shape(uo, f) || mode != MGP_MODE_CONSISTENT)
        return MGP_ERR_ARG;
    const Box bx = box_of(b);
    for (int d = 0; d < 3; d++)
        if (b[2 * d] < 0 || b[2 * d + 1] > uo->shape[d]) return MGP_ERR_ARG;
    if ((bx.x0 == 0 && faces->type[MGP_WEST] != MGP_NEUMANN) ||
        (bx.x1 == uo->shape[0] && faces->type[MGP_EAST] != MGP_NEUMANN) ||
        (bx.y0 == 0 && faces->type[MGP_SOUTH] != MGP_NEUMANN) ||
        (bx.y1 == uo->shape[1] && faces->type[MGP_NORTH] != MGP_NEUMANN) ||
        (bx.z0 == 0 && faces->type[MGP_BOTTOM] != MGP_NEUMANN) ||
        (bx.z1 == uo->shape[2] && faces->type[MGP_TOP] != MGP_NEUMANN))
        return MGP_ERR_ARG;
    for (int q = 0; q < 6; q++)
        if (faces->type[q] == MGP_NEUMANN && !faces->g[q]) return MGP_ERR_ARG;
    if (box_empty(bx)) return MGP_ERR_ARG;
    return march_sweep2<T>(mode, uo, uo, f, faces, b, omega, stream, nullptr, true);
}

// LDS bytes the multi-sweep kernel needs for a level, 0 when it does not fit
static inline size_t lds_sweeps_bytes(const mgp_array* u, size_t esize) {
    const size_t b = ((size_t)(u->shape[0] + 2) * (u->shape[1] + 2) * (u->shape[2] + 2) +
                      (size_t)u->shape[0] * u->shape[1] * u->shape[2]) * esize;
    return b <= 65536 ? b : 0;
}

template <class S, class T = S>
static int rbgs_sweeps(int mode, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* faces,
                       const int64_t b[6], double omega, int64_t nsweeps, void* stream) {
    if (!valid_array(ui) || !valid_array(uo) || !valid_array(f) || !faces || !same_shape(ui, uo) ||
        !same_shape(ui, f) || nsweeps < 0 || (mode != MGP_MODE_REFERENCE && mode != MGP_MODE_CONSISTENT))
        return MGP_ERR_ARG;
    if (!checkerboard(mode, ui) || lin0_of(ui) != lin0_of(uo) || ui->stride[0] != uo->stride[0] ||
        ui->stride[1] != uo->stride[1])
        return MGP_ERR_ARG;
    for (int d = 0; d < 3; d++)
        if (b[2 * d] < 0 || b[2 * d + 1] > ui->shape[d]) return MGP_ERR_ARG;
    const Box bx = box_of(b);
    if ((bx.x0 == 0 && faces->type[MGP_WEST] != MGP_NEUMANN) ||
        (bx.x1 == ui->shape[0] && faces->type[MGP_EAST] != MGP_NEUMANN) ||
        (bx.y0 == 0 && faces->type[MGP_SOUTH] != MGP_NEUMANN) ||
        (bx.y1 == ui->shape[1] && faces->type[MGP_NORTH] != MGP_NEUMANN) ||
        (bx.z0 == 0 && faces->type[MGP_BOTTOM] != MGP_NEUMANN) ||
        (bx.z1 == ui->shape[2] && faces->type[MGP_TOP] != MGP_NEUMANN))
        return MGP_ERR_ARG;
    for (int q = 0; q < 6; q++)
        if (faces->type[q] == MGP_NEUMANN && !faces->g[q]) return MGP_ERR_ARG;
    const size_t lds = lds_sweeps_bytes(ui, sizeof(T));
    if (!lds) return MGP_ERR_ARG;  // too large for one workgroup
    if ((b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]) > 8 * 1024) return MGP_ERR_ARG;  // 8 points per thread
    const int64_t n0 = ui->shape[0], n1 = ui->shape[1], n2 = ui->shape[2];
    if (n1 * n2 > 1024 || n0 * n2 > 1024 || n0 * n1 > 1024) return MGP_ERR_ARG;  // one ghost per thread and face
    if (box_empty(bx) || nsweeps == 0) return ui->data == uo->data ? MGP_OK : MGP_ERR_ARG;
    const RelaxConsts<T> c(omega, faces->h);
    const Faces<S> F = faces_of<S>(*faces);
    const T two_h = (T)(2.0 * faces->h);  // include/neumann.h:108, in the arithmetic type
    const View<S> vi = view_of<S>(*ui);
    const hipStream_t s = (hipStream_t)stream;
    if (mode == MGP_MODE_REFERENCE)
        hipLaunchKernelGGL((k_rbgs_lds<S, T, MGP_MODE_REFERENCE>), dim3(1), dim3(1024), lds, s, vi, view_of<S>(*uo),
                           view_of<S>(*f), F, bx, nsweeps, c.one_omega, c.omega_sixth, c.hsq, two_h, lin0_of(ui));
    else
        hipLaunchKernelGGL((k_rbgs_lds<S, T, MGP_MODE_CONSISTENT>), dim3(1), dim3(1024), lds, s, vi, view_of<S>(*uo),
                           view_of<S>(*f), F, bx, nsweeps, c.one_omega, c.omega_sixth, c.hsq, two_h, lin0_of(ui));
    return launch_status();
}

// the coarsest-level CG solve (consistent mode; see k_cg_lds)
template <class S>
static int cg_solve(const mgp_array* u, const mgp_array* f, const mgp_faces* faces, const int64_t b[6],
                    int64_t max_iters, double rtol, void* stream) {
    if (!valid_array(u) || !valid_array(f) || !faces || !same_shape(u, f) || max_iters < 0 || !(rtol >= 0.0) ||
        !(faces->h > 0.0))
        return MGP_ERR_ARG;
    for (int d = 0; d < 3; d++)
        if (b[2 * d] < 0 || b[2 * d + 1] > u->shape[d]) return MGP_ERR_ARG;
    const Box bx = box_of(b);
    if ((bx.x0 == 0 && faces->type[MGP_WEST] != MGP_NEUMANN) ||
        (bx.x1 == u->shape[0] && faces->type[MGP_EAST] != MGP_NEUMANN) ||
        (bx.y0 == 0 && faces->type[MGP_SOUTH] != MGP_NEUMANN) ||
        (bx.y1 == u->shape[1] && faces->type[MGP_NORTH] != MGP_NEUMANN) ||
        (bx.z0 == 0 && faces->type[MGP_BOTTOM] != MGP_NEUMANN) ||
        (bx.z1 == u->shape[2] && faces->type[MGP_TOP] != MGP_NEUMANN))
        return MGP_ERR_ARG;
    // a Neumann face must be the box's face (its points are unknowns)
    if ((faces->type[MGP_WEST] == MGP_NEUMANN && bx.x0 != 0) ||
        (faces->type[MGP_EAST] == MGP_NEUMANN && bx.x1 != u->shape[0]) ||
        (faces->type[MGP_SOUTH] == MGP_NEUMANN && bx.y0 != 0) ||
        (faces->type[MGP_NORTH] == MGP_NEUMANN && bx.y1 != u->shape[1]) ||
        (faces->type[MGP_BOTTOM] == MGP_NEUMANN && bx.z0 != 0) ||
        (faces->type[MGP_TOP] == MGP_NEUMANN && bx.z1 != u->shape[2]))
        return MGP_ERR_ARG;
    const int64_t n0 = u->shape[0], n1 = u->shape[1], n2 = u->shape[2];
    const size_t lds = (size_t)(n0 + 2) * (n1 + 2) * (n2 + 2) * sizeof(double);
    if (lds > 65536 - 256) return MGP_ERR_ARG;  // with the kernel's 128-byte reduction array
    if ((b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]) > 8 * 1024) return MGP_ERR_ARG;
    if (n1 * n2 > 4096 || n0 * n2 > 4096 || n0 * n1 > 4096) return MGP_ERR_ARG;  // 4 ghost cells per thread
    if (box_empty(bx) || max_iters == 0) return MGP_OK;
    const Faces<S> F = faces_of<S>(*faces);
    hipLaunchKernelGGL(k_cg_lds<S>, dim3(1), dim3(1024), lds, (hipStream_t)stream, view_of<S>(*u), view_of<S>(*f), F,
                       bx, max_iters, rtol, 2.0 * faces->h, 1.0 / (faces->h * faces->h));
    return launch_status();
}

template <class T>
static int jacobi(int mode, const mgp_array* u, const mgp_array* v, const mgp_array* f, const int64_t b[6],
                  double omega, double h, void* stream) {
    if (!valid_array(u) || !valid_array(v) || !valid_array(f) || !same_shape(u, v) || !same_shape(u, f) ||
        !valid_box(v, b))
        return MGP_ERR_ARG;
    const Box bx = box_of(b);
    if (box_empty(bx)) return MGP_OK;
    const BoxGrid g = box_grid(bx);
    const RelaxConsts<T> c(omega, h);
    if (mode == MGP_MODE_REFERENCE)
        hipLaunchKernelGGL((k_jacobi<T, MGP_MODE_REFERENCE>), g.grid, g.block, 0, (hipStream_t)stream, view_of<T>(*u),
                           view_of<T>(*v), view_of<T>(*f), bx, g.zchunks, c.one_omega, c.omega_sixth, c.hsq);
    else if (mode == MGP_MODE_CONSISTENT)
        hipLaunchKernelGGL((k_jacobi<T, MGP_MODE_CONSISTENT>), g.grid, g.block, 0, (hipStream_t)stream, view_of<T>(*u),
                           view_of<T>(*v), view_of<T>(*f), bx, g.zchunks, c.one_omega, c.omega_sixth, c.hsq);
    else
        return MGP_ERR_ARG;
    return launch_status();
}

template <class T>
static int residual(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t b[6], double h,
                    double* norm, void* stream) {
    if (!valid_array(u) || !valid_array(f) || !valid_array(r) || !same_shape(u, f) || !same_shape(u, r) ||
        !valid_box(u, b) || (mode != MGP_MODE_REFERENCE && mode != MGP_MODE_CONSISTENT))
        return MGP_ERR_ARG;
    const Box bx = box_of(b);
    if (g_resid_kernel == 2 && march_residual<T>(mode, u, f, r, b, h, norm, stream) == MGP_OK) return MGP_OK;
    const T hsq = (T)(h * h);  // include/residual.h:23
    const T inv_h2 = (T)(1.0 / (h * h));
    const hipStream_t s = (hipStream_t)stream;
    if (norm) {
        const Box sh = shape_box(r);
        const BoxGrid g = column_grid(sh);
        if (mode == MGP_MODE_REFERENCE)
            hipLaunchKernelGGL((k_residual_norm<T, MGP_MODE_REFERENCE>), g.grid, g.block, 0, s, view_of<T>(*u),
                               view_of<T>(*f), view_of<T>(*r), sh, bx, hsq, inv_h2, (unsigned long long*)norm);
        else
            hipLaunchKernelGGL((k_residual_norm<T, MGP_MODE_CONSISTENT>), g.grid, g.block, 0, s, view_of<T>(*u),
                               view_of<T>(*f), view_of<T>(*r), sh, bx, hsq, inv_h2, (unsigned long long*)norm);
    } else {
        if (box_empty(bx)) return MGP_OK;
        const BoxGrid g = box_grid(bx);
        if (mode == MGP_MODE_REFERENCE)
            hipLaunchKernelGGL((k_residual<T, MGP_MODE_REFERENCE>), g.grid, g.block, 0, s, view_of<T>(*u),
                               view_of<T>(*f), view_of<T>(*r), bx, g.zchunks, hsq, inv_h2);
        else
            hipLaunchKernelGGL((k_residual<T, MGP_MODE_CONSISTENT>), g.grid, g.block, 0, s, view_of<T>(*u),
                               view_of<T>(*f), view_of<T>(*r), bx, g.zchunks, hsq, inv_h2);
    }
    return launch_status();
}

static inline bool coarse_of(const mgp_array* fine, const mgp_array* coarse) {
    for (int d = 0; d < 3; d++)
        if ((fine->shape[d] - 1) / 2 + 1 != coarse->shape[d]) return false;
    return true;
}

template <class T>
static int restrict_any(bool fw, const mgp_array* in, const mgp_array* out, void* stream) {
    if (!valid_array(in) || !valid_array(out) || !coarse_of(in, out)) return MGP_ERR_ARG;
    const Box bx = shape_box(out);
    const BoxGrid g = box_grid(bx);
    if (fw)
        hipLaunchKernelGGL(k_fw<T>, g.grid, g.block, 0, (hipStream_t)stream, view_of<T>(*in), view_of<T>(*out), bx,
                           g.zchunks);
    else
        hipLaunchKernelGGL(k_inject<T>, g.grid, g.block, 0, (hipStream_t)stream, view_of<T>(*in), view_of<T>(*out), bx,
                           g.zchunks);
    return launch_status();
}

template <class T, bool ADD>
static int prolong(const mgp_array* coarse, const mgp_array* fine, void* stream) {
    if (!valid_array(coarse) || !valid_array(fine)) return MGP_ERR_ARG;
    for (int d = 0; d < 3; d++)  // exact 2:1 nesting (trilinearinterpolation.h:31-37)
        if (2 * (coarse->shape[d] - 1) + 1 != fine->shape[d]) return MGP_ERR_ARG;
    if (ADD && g_prolong_kernel == 2 && march_prolong_add<T>(coarse, fine, stream) == MGP_OK) return MGP_OK;
    const Box bx = shape_box(fine);
    const BoxGrid g = box_grid(bx);
    hipLaunchKernelGGL((k_prolong<T, ADD>), g.grid, g.block, 0, (hipStream_t)stream, view_of<T>(*coarse),
                       view_of<T>(*fine), bx, g.zchunks);
    return launch_status();
}

static inline int64_t face_points(int loc, const mgp_array* u) {
    int64_t n = 1;
    for (int d = 0; d < 3; d++) {
        const bool flat = (d == 0 && (loc == MGP_EAST || loc == MGP_WEST)) ||
                          (d == 1 && (loc == MGP_NORTH || loc == MGP_SOUTH)) ||
                          (d == 2 && (loc == MGP_TOP || loc == MGP_BOTTOM));
        n *= flat ? 1 : u->shape[d];
    }
    return n;
}

static inline int64_t halo_of_face(const mgp_array* u, int loc) { return u->halo[loc]; }

template <class T>
static int bc_dirichlet(const mgp_array* u, int loc, const void* g, void* stream) {
    if (!valid_array(u) || loc < 0 || loc > 5 || !g) return MGP_ERR_ARG;
    const int64_t n = face_points(loc, u);
    hipLaunchKernelGGL(k_bc_dirichlet<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       view_of<T>(*u), loc, (const T*)g);
    return launch_status();
}

template <class T>
static int bc_neumann(const mgp_array* u, int loc, const void* g, double h, void* stream) {
    if (!valid_array(u) || loc < 0 || loc > 5 || halo_of_face(u, loc) < 1) return MGP_ERR_ARG;
    for (int d = 0; d < 3; d++)
        if (u->shape[d] < 3) return MGP_ERR_ARG;  // the mirror point must exist
    const int64_t n = face_points(loc, u);
    const T two_h = (T)(2.0 * h);  // include/neumann.h:108
    hipLaunchKernelGGL(k_bc_neumann<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       view_of<T>(*u), loc, (const T*)g, two_h);
    return launch_status();
}

template <class T>
static int bc_neumann_faces(const mgp_array* u, const mgp_faces* faces, void* stream) {
    if (!valid_array(u) || !faces) return MGP_ERR_ARG;
    NeumannFaces<T> F{};
    F.start[0] = 0;
    for (int q = 0; q < 6; q++) {
        const bool nm = faces->type[q] == MGP_NEUMANN;
        if (nm && halo_of_face(u, q) < 1) return MGP_ERR_ARG;
        F.g[q] = nm ? static_cast<const T*>(faces->g[q]) : nullptr;
        F.start[q + 1] = F.start[q] + (nm ? face_points(q, u) : 0);
    }
    if (F.start[6] == 0) return MGP_OK;
    for (int d = 0; d < 3; d++)
        if (u->shape[d] < 3) return MGP_ERR_ARG;  // the mirror point must exist
    F.two_h = (T)(2.0 * faces->h);  // include/neumann.h:108
    hipLaunchKernelGGL(k_bc_neumann_faces<T>, dim3((unsigned)((F.start[6] + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, view_of<T>(*u), F);
    return launch_status();
}

static inline mgp_array face_array(void* data, int loc, const int64_t n[3]) {
    mgp_array a{};
    a.data = data;
    for (int d = 0; d < 3; d++) a.shape[d] = n[d];
    if (loc == MGP_EAST || loc == MGP_WEST) a.shape[0] = 1;
    if (loc == MGP_NORTH || loc == MGP_SOUTH) a.shape[1] = 1;
    if (loc == MGP_TOP || loc == MGP_BOTTOM) a.shape[2] = 1;
    a.stride[2] = 1;
    a.stride[1] = a.shape[2];
    a.stride[0] = a.shape[1] * a.shape[2];
    return a;
}

template <class T>
static int bc_restrict_neumann(int restriction, const mgp_array* uf, int loc, const void* gf, void* gc, double h_fine,
                               void* stream) {
    if (!valid_array(uf) || loc < 0 || loc > 5 || !gf || !gc ||
        (restriction != MGP_INJECTION && restriction != MGP_FULL_WEIGHTING))
        return MGP_ERR_ARG;
    for (int d = 0; d < 3; d++)
        if (uf->shape[d] < 3 || (uf->shape[d] & 1) == 0) return MGP_ERR_ARG;
    const int64_t nc[3] = {uf->shape[0] / 2 + 1, uf->shape[1] / 2 + 1, uf->shape[2] / 2 + 1};
    const mgp_array F = face_array((void*)gf, loc, uf->shape);
    const mgp_array C = face_array(gc, loc, nc);
    // neumann.h:198-211
    T c1 = -3.0 / 2.0, c2 = 2.0, c3 = -0.5;
    if (loc == MGP_EAST || loc == MGP_NORTH || loc == MGP_TOP) {
        c3 = 0.5;
        c2 = -2.0;
        c1 = 3.0 / 2.0;
    }
    const T hf = (T)h_fine;  // neumann.h:221
    const int64_t n = C.shape[0] * C.shape[1] * C.shape[2];
    hipLaunchKernelGGL(k_bc_restrict_neumann<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       restriction, view_of<T>(*uf), loc, view_of<T>(F), view_of<T>(C), c1, c2, c3, hf);
    return launch_status();
}

template <class T>
static int residual_restrict_fw(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r,
                                const mgp_array* cf, const int64_t b[6], double h, void* stream) {
    if (mode != MGP_MODE_CONSISTENT || !valid_array(u) || !valid_array(f) || !valid_array(r) || !valid_array(cf) ||
        !same_shape(u, f) || !same_shape(u, r) || !coarse_of(u, cf) || !valid_box(u, b))
        return MGP_ERR_ARG;
    return march_residual_fw<T>(u, f, r, cf, b, h, stream);
}


// ---- FP16 storage: validation and launches
static inline HalfView half_view(const mgp_array* a) { return HalfView(view_of<_Float16>(*a)); }

// box strictly inside the shape (all faces Dirichlet)
static inline bool interior_box(const mgp_array* a, const int64_t b[6]) {
    for (int d = 0; d < 3; d++)
        if (b[2 * d] < b[2 * d + 1] && (b[2 * d] < 1 || b[2 * d + 1] > a->shape[d] - 1)) return false;
    return true;
}

static int fill_f16(const mgp_array* a, double value, void* stream) {
    if (!a || !a->data) return MGP_ERR_ARG;
    const int64_t n = a->stride[0] * (a->shape[0] + a->halo[MGP_EAST] + a->halo[MGP_WEST]);
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_h, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (_Float16*)a->data, n,
                       (float)value);
    return launch_status();
}

static int rbgs_half_f16(const mgp_array* u, const mgp_array* f, const int64_t b[6], double omega, double h,
                         int colour, void* stream) {
    if (!valid_array(u) || !valid_array(f) || !same_shape(u, f) || !interior_box(u, b) || (colour & ~1))
        return MGP_ERR_ARG;
    const Box bx = box_of(b);
    if (box_empty(bx)) return MGP_OK;
    const BoxGrid g = box_grid(bx);
    const RelaxConsts<float> c(omega, h);
    hipLaunchKernelGGL(k_gs_half_h, g.grid, g.block, 0, (hipStream_t)stream, half_view(u), half_view(f), bx, g.zchunks,
                       c.omega_sixth, c.hsq, colour);
    return launch_status();
}

static int rbgs_sweeps_f16(const mgp_array* u, const mgp_array* f, const int64_t b[6], double omega, double h,
                           int64_t nsweeps, void* stream) {
    if (!valid_array(u) || !valid_array(f) || !same_shape(u, f) || !interior_box(u, b) || nsweeps < 0)
        return MGP_ERR_ARG;
    const int64_t N = u->shape[0] * u->shape[1] * u->shape[2];
    if (N > 8192) return MGP_ERR_ARG;  // u and f as float in 64 KiB of LDS, 8 box points per thread
    const Box bx = box_of(b);
    if (box_empty(bx) || nsweeps == 0) return MGP_OK;
    const RelaxConsts<float> c(omega, h);
    hipLaunchKernelGGL(k_sweeps_lds_h, dim3(1), dim3(1024), (size_t)(2 * N * sizeof(float)), (hipStream_t)stream,
                       half_view(u), half_view(f), bx, nsweeps, c.omega_sixth, c.hsq);
    return launch_status();
}

static int residual_f16(const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t b[6], double h,
                        double* norm, void* stream) {
    if (!valid_array(u) || !valid_array(f) || !valid_array(r) || !same_shape(u, f) || !same_shape(u, r) ||
        !interior_box(u, b))
        return MGP_ERR_ARG;
    const Box bx = box_of(b), sh = shape_box(r);
    const BoxGrid g = column_grid(sh);
    const float inv_h2 = (float)(1.0 / (h * h));  // as the FP32 residual forms it
    hipLaunchKernelGGL(k_residual_norm_h, g.grid, g.block, 0, (hipStream_t)stream, half_view(u), half_view(f),
                       half_view(r), sh, bx, inv_h2, (unsigned long long*)norm);
    return launch_status();
}

static int restrict_fw_scaled(bool from_f32, const mgp_array* in, const mgp_array* out, const double* norm,
                              void* stream) {
    if (!valid_array(in) || !valid_array(out) || !coarse_of(in, out) || !norm) return MGP_ERR_ARG;
    const Box bx = shape_box(out);
    const BoxGrid g = box_grid(bx);
    const hipStream_t s = (hipStream_t)stream;
    if (from_f32)
        hipLaunchKernelGGL(k_fw_scaled_h<View<float>>, g.grid, g.block, 0, s, view_of<float>(*in), half_view(out), bx,
                           g.zchunks, norm);
    else
        hipLaunchKernelGGL(k_fw_scaled_h<HalfView>, g.grid, g.block, 0, s, half_view(in), half_view(out), bx,
                           g.zchunks, norm);
    return launch_status();
}

static int prolong_add_scaled(bool to_f32, const mgp_array* coarse, const mgp_array* fine, const double* norm,
                              void* stream) {
    if (!valid_array(coarse) || !valid_array(fine) || !norm) return MGP_ERR_ARG;
    for (int d = 0; d < 3; d++)
        if (2 * (coarse->shape[d] - 1) + 1 != fine->shape[d]) return MGP_ERR_ARG;
    const Box bx = shape_box(fine);
    const BoxGrid g = box_grid(bx);
    const hipStream_t s = (hipStream_t)stream;
    if (to_f32)
        hipLaunchKernelGGL(k_prolong_add_scaled_h<View<float>>, g.grid, g.block, 0, s, half_view(coarse),
                           view_of<float>(*fine), bx, g.zchunks, norm);
    else
        hipLaunchKernelGGL(k_prolong_add_scaled_h<HalfView>, g.grid, g.block, 0, s, half_view(coarse),
                           half_view(fine), bx, g.zchunks, norm);
    return launch_status();
}

}  // namespace mgp

// ============================================================ C entry points
using namespace mgp;

extern "C" {

int mgp_fill_f64(const mgp_array* a, double v, void* s) { return fill<double>(a, v, s); }
int mgp_fill_f32(const mgp_array* a, double v, void* s) { return fill<float>(a, v, s); }
int mgp_add_f64(const mgp_array* a, const mgp_array* b, void* s) { return add<double>(a, b, s); }
int mgp_add_f32(const mgp_array* a, const mgp_array* b, void* s) { return add<float>(a, b, s); }
int mgp_norm_inf_f64(const mgp_array* a, double* o, void* s) { return norm_inf<double>(a, o, s); }
int mgp_norm_inf_f32(const mgp_array* a, double* o, void* s) { return norm_inf<float>(a, o, s); }

int mgp_cast_f64_f32(const mgp_array* in, const mgp_array* out, void* s) {
    if (!valid_array(in) || !valid_array(out) || !same_shape(in, out)) return MGP_ERR_ARG;
    const Box bx = shape_box(in);
    const BoxGrid g = box_grid(bx);
    hipLaunchKernelGGL((k_cast<double, float>), g.grid, g.block, 0, (hipStream_t)s, view_of<double>(*in),
                       view_of<float>(*out), bx, g.zchunks);
    return launch_status();
}

int mgp_cast_f32_f64(const mgp_array* in, const mgp_array* out, void* s) {
    if (!valid_array(in) || !valid_array(out) || !same_shape(in, out)) return MGP_ERR_ARG;
    const Box bx = shape_box(in);
    const BoxGrid g = box_grid(bx);
    hipLaunchKernelGGL((k_cast<float, double>), g.grid, g.block, 0, (hipStream_t)s, view_of<float>(*in),
                       view_of<double>(*out), bx, g.zchunks);
    return launch_status();
}

int mgp_update_residual_f64_f32(int m, const mgp_array* u, const mgp_array* u_new, const mgp_array* e,
                                const mgp_array* f, const mgp_array* r, const mgp_array* out, const mgp_faces* faces,
                                const int64_t b[6], double* norm, void* s) {
    if (!valid_array(u) || !valid_array(u_new) || !valid_array(e) || !valid_array(f) || !valid_array(r) ||
        !valid_array(out) || !faces || !same_shape(u, u_new) || !same_shape(u, e) || !same_shape(u, f) ||
        !same_shape(u, r) || !same_shape(u, out) || u->data == u_new->data || !valid_box(u, b) ||
        (m != MGP_MODE_REFERENCE && m != MGP_MODE_CONSISTENT) || !(faces->h > 0.0))
        return MGP_ERR_ARG;
    // a box point on a face needs that face's ghost: Neumann with its data
    const Box bx = box_of(b);
    if ((bx.x0 == 0 && faces->type[MGP_WEST] != MGP_NEUMANN) ||
        (bx.x1 == u->shape[0] && faces->type[MGP_EAST] != MGP_NEUMANN) ||
        (bx.y0 == 0 && faces->type[MGP_SOUTH] != MGP_NEUMANN) ||
        (bx.y1 == u->shape[1] && faces->type[MGP_NORTH] != MGP_NEUMANN) ||
        (bx.z0 == 0 && faces->type[MGP_BOTTOM] != MGP_NEUMANN) ||
        (bx.z1 == u->shape[2] && faces->type[MGP_TOP] != MGP_NEUMANN))
        return MGP_ERR_ARG;
    for (int q = 0; q < 6; q++)
        if (faces->type[q] == MGP_NEUMANN && !faces->g[q]) return MGP_ERR_ARG;
    return march_update_residual(m, u, u_new, e, f, r, out, faces, b, norm, s);
}

// box and faces of the hi + lo kernels: a box point on a face needs that
// face's ghost (Neumann with its data)
static bool hilo_faces_ok(const mgp_array* u, const mgp_faces* faces, const int64_t b[6]) {
    if (!faces || !(faces->h > 0.0) || !valid_box(u, b)) return false;
    const Box bx = box_of(b);
    if ((bx.x0 == 0 && faces->type[MGP_WEST] != MGP_NEUMANN) ||
        (bx.x1 == u->shape[0] && faces->type[MGP_EAST] != MGP_NEUMANN) ||
        (bx.y0 == 0 && faces->type[MGP_SOUTH] != MGP_NEUMANN) ||
        (bx.y1 == u->shape[1] && faces->type[MGP_NORTH] != MGP_NEUMANN) ||
        (bx.z0 == 0 && faces->type[MGP_BOTTOM] != MGP_NEUMANN) ||
        (bx.z1 == u->shape[2] && faces->type[MGP_TOP] != MGP_NEUMANN))
        return false;
    for (int q = 0; q < 6; q++)
        if (faces->type[q] == MGP_NEUMANN && !faces->g[q]) return false;
    return true;
}

int mgp_update_residual_hilo_f64_f32(const mgp_array* u, const mgp_array* u_new, const mgp_array* lo,
                                     const mgp_array* lo_new, const mgp_array* e, const mgp_array* f,
                                     const mgp_array* r, const mgp_array* out, const mgp_faces* faces,
                                     const int64_t b[6], double* norm, void* s) {
    const mgp_array* all[] = {u, u_new, lo, lo_new, e, f, r, out};
    for (const mgp_array* a : all)
        if (!valid_array(a) || !same_shape(u, a)) return MGP_ERR_ARG;
    if (u->data == u_new->data || lo->data == lo_new->data || !hilo_faces_ok(u, faces, b)) return MGP_ERR_ARG;
    return march_hilo_residual(u, u_new, lo, lo_new, e, f, r, out, faces, b, norm, s);
}

int mgp_residual_hilo_f64(const mgp_array* u, const mgp_array* lo, const mgp_array* f, const mgp_array* r,
                          const mgp_faces* faces, const int64_t b[6], double* norm, void* s) {
    const mgp_array* all[] = {u, lo, f, r};
    for (const mgp_array* a : all)
        if (!valid_array(a) || !same_shape(u, a)) return MGP_ERR_ARG;
    if (r->data == u->data || !hilo_faces_ok(u, faces, b)) return MGP_ERR_ARG;
    return march_hilo_residual(u, nullptr, lo, nullptr, nullptr, f, r, r, faces, b, norm, s);
}

int mgp_add_f32_to_f64(const mgp_array* u, const mgp_array* e, void* s) {
    if (!valid_array(u) || !valid_array(e) || !same_shape(u, e)) return MGP_ERR_ARG;
    const Box bx = shape_box(u);
    const BoxGrid g = box_grid(bx);
    hipLaunchKernelGGL(k_add_f32_to_f64, g.grid, g.block, 0, (hipStream_t)s, view_of<double>(*u), view_of<float>(*e),
                       bx, g.zchunks);
    return launch_status();
}

int mgp_fill_modes_f64(const mgp_array* f, int64_t nmodes, const int32_t* abc, const double* amp, double h,
                       void* s) {
    if (!valid_array(f) || nmodes < 0 || nmodes > MGP_MAX_MODES || (nmodes && (!abc || !amp)) || !(h > 0.0))
        return MGP_ERR_ARG;
    Modes m;
    m.n = (int32_t)nmodes;
    for (int64_t q = 0; q < nmodes; q++) {
        for (int d = 0; d < 3; d++) m.abc[3 * q + d] = abc[3 * q + d];
        m.amp[q] = amp[q];
    }
    const Box bx = shape_box(f);
    const BoxGrid g = box_grid(bx);
    hipLaunchKernelGGL(k_fill_modes<double>, g.grid, g.block, 0, (hipStream_t)s, view_of<double>(*f), bx, g.zchunks,
                       m, h);
    return launch_status();
}

int mgp_residual_restrict_fw_f64(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r,
                                 const mgp_array* cf, const int64_t b[6], double h, void* s) {
    return residual_restrict_fw<double>(m, u, f, r, cf, b, h, s);
}
int mgp_residual_restrict_fw_f32(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r,
                                 const mgp_array* cf, const int64_t b[6], double h, void* s) {
    return residual_restrict_fw<float>(m, u, f, r, cf, b, h, s);
}

int mgp_residual_norm_f64(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t b[6],
                          double h, double* norm, void* s) {
    if (!valid_array(u) || !valid_array(f) || !valid_array(r) || !norm || !same_shape(u, f) || !same_shape(u, r) ||
        !valid_box(u, b) || (m != MGP_MODE_REFERENCE && m != MGP_MODE_CONSISTENT))
        return MGP_ERR_ARG;
    return march_residual_to<double, double>(m, u, f, r, r, b, h, norm, s, false);
}
int mgp_residual_norm_f32(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t b[6],
                          double h, double* norm, void* s) {
    if (!valid_array(u) || !valid_array(f) || !valid_array(r) || !norm || !same_shape(u, f) || !same_shape(u, r) ||
        !valid_box(u, b) || (m != MGP_MODE_REFERENCE && m != MGP_MODE_CONSISTENT))
        return MGP_ERR_ARG;
    return march_residual_to<float, float>(m, u, f, r, r, b, h, norm, s, false);
}

int mgp_residual_cast_f64_f32(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r,
                              const mgp_array* out, const int64_t b[6], double h, double* norm, void* s) {
    if (!valid_array(u) || !valid_array(f) || !valid_array(r) || !valid_array(out) || !same_shape(u, f) ||
        !same_shape(u, r) || !same_shape(u, out) || !valid_box(u, b) ||
        (m != MGP_MODE_REFERENCE && m != MGP_MODE_CONSISTENT))
        return MGP_ERR_ARG;
    return march_residual_to<double, float>(m, u, f, r, out, b, h, norm, s);
}

int mgp_pack_f64(const mgp_array* a, void* out, void* s) {
    if (!valid_array(a) || !out) return MGP_ERR_ARG;
    const Box bx = shape_box(a);
    const BoxGrid g = box_grid(bx);
    hipLaunchKernelGGL(k_pack<double>, g.grid, g.block, 0, (hipStream_t)s, view_of<double>(*a), bx, g.zchunks,
                       static_cast<double*>(out));
    return launch_status();
}

int mgp_zero_scalars(double* p, int64_t n, void* s) {
    if (!p || n < 0) return MGP_ERR_ARG;
    if (n == 0) return MGP_OK;
    hipLaunchKernelGGL(k_zero_scalars, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s, p, n);
    return launch_status();
}

int mgp_rbgs_half_f64(int m, const mgp_array* u, const mgp_array* f, const int64_t b[6], double w, double h, int c,
                      void* s) {
    return rbgs_half<double>(m, u, f, b, w, h, c, s);
}
int mgp_rbgs_half_f32(int m, const mgp_array* u, const mgp_array* f, const int64_t b[6], double w, double h, int c,
                      void* s) {
    return rbgs_half<float>(m, u, f, b, w, h, c, s);
}
int mgp_rbgs_sweep_f64(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,
                       const int64_t b[6], double w, void* s) {
    return rbgs_sweep<double>(m, ui, uo, f, F, b, w, s);
}
int mgp_rbgs_sweep_f32(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,
                       const int64_t b[6], double w, void* s) {
    return rbgs_sweep<float>(m, ui, uo, f, F, b, w, s);
}
int mgp_rbgs_sweep2_f64(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,
                        const int64_t b[6], double w, void* s) {
    return rbgs_sweep2<double>(m, ui, uo, f, F, b, w, s);
}
int mgp_rbgs_sweep2_f32(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,
                        const int64_t b[6], double w, void* s) {
    return rbgs_sweep2<float>(m, ui, uo, f, F, b, w, s);
}
int mgp_rbgs_sweep2_from_zero_f64(int m, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,
                                  const int64_t b[6], double w, void* s) {
    return rbgs_sweep2_from_zero<double>(m, uo, f, F, b, w, s);
}
int mgp_rbgs_sweep2_from_zero_f32(int m, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,
                                  const int64_t b[6], double w, void* s) {
    return rbgs_sweep2_from_zero<float>(m, uo, f, F, b, w, s);
}
int mgp_rbgs_sweep2_prolong_f64(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f,
                                const mgp_faces* F, const int64_t b[6], double w, const mgp_array* c, void* s) {
    if (!c) return MGP_ERR_ARG;
    return rbgs_sweep2<double>(m, ui, uo, f, F, b, w, s, c);
}
int mgp_rbgs_sweep2_prolong_f32(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f,
                                const mgp_faces* F, const int64_t b[6], double w, const mgp_array* c, void* s) {
    if (!c) return MGP_ERR_ARG;
    return rbgs_sweep2<float>(m, ui, uo, f, F, b, w, s, c);
}
int mgp_rbgs_sweeps_f64(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,
                        const int64_t b[6], double w, int64_t n, void* s) {
    return rbgs_sweeps<double>(m, ui, uo, f, F, b, w, n, s);
}
int mgp_rbgs_sweeps_f32(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,
                        const int64_t b[6], double w, int64_t n, void* s) {
    return rbgs_sweeps<float>(m, ui, uo, f, F, b, w, n, s);
}
int mgp_rbgs_sweeps_f32_f64(int m, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* F,
                            const int64_t b[6], double w, int64_t n, void* s) {
    return rbgs_sweeps<float, double>(m, ui, uo, f, F, b, w, n, s);
}
int mgp_cg_solve_f64(const mgp_array* u, const mgp_array* f, const mgp_faces* F, const int64_t b[6], int64_t max_iters,
                     double rtol, void* s) {
    return cg_solve<double>(u, f, F, b, max_iters, rtol, s);
}
int mgp_cg_solve_f32(const mgp_array* u, const mgp_array* f, const mgp_faces* F, const int64_t b[6], int64_t max_iters,
                     double rtol, void* s) {
    return cg_solve<float>(u, f, F, b, max_iters, rtol, s);
}
int mgp_jacobi_f64(int m, const mgp_array* u, const mgp_array* v, const mgp_array* f, const int64_t b[6], double w,
                   double h, void* s) {
    return jacobi<double>(m, u, v, f, b, w, h, s);
}
int mgp_jacobi_f32(int m, const mgp_array* u, const mgp_array* v, const mgp_array* f, const int64_t b[6], double w,
                   double h, void* s) {
    return jacobi<float>(m, u, v, f, b, w, h, s);
}
int mgp_residual_f64(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t b[6], double h,
                     double* n, void* s) {
    return residual<double>(m, u, f, r, b, h, n, s);
}
int mgp_residual_f32(int m, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t b[6], double h,
                     double* n, void* s) {
    return residual<float>(m, u, f, r, b, h, n, s);
}
int mgp_restrict_inject_f64(const mgp_array* i, const mgp_array* o, void* s) { return restrict_any<double>(false, i, o, s); }
int mgp_restrict_inject_f32(const mgp_array* i, const mgp_array* o, void* s) { return restrict_any<float>(false, i, o, s); }
int mgp_restrict_fw_f64(const mgp_array* i, const mgp_array* o, void* s) { return restrict_any<double>(true, i, o, s); }
int mgp_restrict_fw_f32(const mgp_array* i, const mgp_array* o, void* s) { return restrict_any<float>(true, i, o, s); }
int mgp_prolong_trilinear_f64(const mgp_array* c, const mgp_array* f, void* s) { return prolong<double, false>(c, f, s); }
int mgp_prolong_trilinear_f32(const mgp_array* c, const mgp_array* f, void* s) { return prolong<float, false>(c, f, s); }
int mgp_prolong_add_f64(const mgp_array* c, const mgp_array* u, void* s) { return prolong<double, true>(c, u, s); }
int mgp_prolong_add_f32(const mgp_array* c, const mgp_array* u, void* s) { return prolong<float, true>(c, u, s); }
int mgp_bc_dirichlet_f64(const mgp_array* u, int l, const void* g, void* s) { return bc_dirichlet<double>(u, l, g, s); }
int mgp_bc_dirichlet_f32(const mgp_array* u, int l, const void* g, void* s) { return bc_dirichlet<float>(u, l, g, s); }
int mgp_bc_neumann_f64(const mgp_array* u, int l, const void* g, double h, void* s) {
    return bc_neumann<double>(u, l, g, h, s);
}
int mgp_bc_neumann_f32(const mgp_array* u, int l, const void* g, double h, void* s) {
    return bc_neumann<float>(u, l, g, h, s);
}
int mgp_bc_neumann_faces_f64(const mgp_array* u, const mgp_faces* faces, void* s) {
    return bc_neumann_faces<double>(u, faces, s);
}
int mgp_bc_neumann_faces_f32(const mgp_array* u, const mgp_faces* faces, void* s) {
    return bc_neumann_faces<float>(u, faces, s);
}
int mgp_bc_restrict_neumann_f64(int r, const mgp_array* u, int l, const void* gf, void* gc, double h, void* s) {
    return bc_restrict_neumann<double>(r, u, l, gf, gc, h, s);
}
int mgp_bc_restrict_neumann_f32(int r, const mgp_array* u, int l, const void* gf, void* gc, double h, void* s) {
    return bc_restrict_neumann<float>(r, u, l, gf, gc, h, s);
}

// FP16 storage of coarse levels (consistent arithmetic, Dirichlet faces)
int mgp_fill_f16(const mgp_array* a, double v, void* s) { return fill_f16(a, v, s); }
int mgp_rbgs_half_f16(const mgp_array* u, const mgp_array* f, const int64_t b[6], double w, double h, int c, void* s) {
    return rbgs_half_f16(u, f, b, w, h, c, s);
}
int mgp_rbgs_sweeps_f16(const mgp_array* u, const mgp_array* f, const int64_t b[6], double w, double h, int64_t n,
                        void* s) {
    return rbgs_sweeps_f16(u, f, b, w, h, n, s);
}
int mgp_residual_f16(const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t b[6], double h,
                     double* norm, void* s) {
    return residual_f16(u, f, r, b, h, norm, s);
}
int mgp_restrict_fw_scaled_f16(const mgp_array* i, const mgp_array* o, const double* norm, void* s) {
    return restrict_fw_scaled(false, i, o, norm, s);
}
int mgp_restrict_fw_scaled_f32_f16(const mgp_array* i, const mgp_array* o, const double* norm, void* s) {
    return restrict_fw_scaled(true, i, o, norm, s);
}
int mgp_prolong_add_scaled_f16(const mgp_array* c, const mgp_array* u, const double* norm, void* s) {
    return prolong_add_scaled(false, c, u, norm, s);
}
int mgp_prolong_add_scaled_f16_f32(const mgp_array* c, const mgp_array* u, const double* norm, void* s) {
    return prolong_add_scaled(true, c, u, norm, s);
}

}  // extern "C"
