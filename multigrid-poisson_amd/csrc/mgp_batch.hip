// Batched multigrid solves (mgp_batch_*): many independent Dirichlet problems
// of uneven shapes, solved together with FP32 V-cycles inside FP64 iterative
// refinement.  No counterpart in the reference, which solves one problem per
// PoissonSolver; each problem here goes through exactly the arithmetic of a
// single RefinementSolver<FullWeighting, TrilinearInterpolation, S> solve in
// the consistent mode (same point functions, same order of operations per
// problem), so a problem's result is the same bits alone or in a batch.
//
// Layout: every level of every problem is a halo-free C-order array in one
// device arena.  A launch covers one operation at one level for every problem
// that takes part in it: a work list of tiles (grid, x-range, y-range, 64
// z-columns) is walked by a persistent grid of 256-thread workgroups, and a
// per-problem active flag lets converged problems drop out without rebuilding
// the lists (so one recorded graph serves every iteration).  Tile extents per
// level shape come from a small autotuner that times the candidates on the
// problem's own arrays and caches the winner per (shape, type).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "mgp.h"
#include "mgp_device.h"
#include "mgp_point.h"

namespace mgp {

// ------------------------------------------------------------ device side

struct BGrid {
    View<float> u[2];  // solution and its Jacobi partner
    View<float> f, r;
    int32_t problem;
    int32_t fine;      // grid index of the next finer level (-1 on level 0)
    float one_omega, omega_sixth, hsq, inv_h2;
};

struct BTop {  // FP64 finest level of a problem (its residual goes straight to the FP32 f)
    View<double> u, f;
    double hsq, inv_h2;
};

struct BItem {
    int32_t grid;
    int32_t x0, x1, y0, y1, z0;
};

struct BState {  // per-problem iteration state (device)
    unsigned long long rnorm;  // max |r| bits of the last residual
    double fnorm;
    double rel;
    int32_t iter;
    int32_t active;
};

enum BOp : int {
    OP_JACOBI = 0,   // u[1-src] = J(u[src]) on the interior
    OP_GS_HALF,      // u[src] red (colour 0) or black points in place
    OP_RESIDUAL,     // r = f - L u[src] on the interior
    OP_FW,           // coarse f = FW(fine r) over the coarse shape (grid = coarse level)
    OP_ZERO_U,       // u[src] = 0 over the shape
    OP_PROLONG_ADD,  // u[src] += P(coarse u[src2]) over the fine shape (grid = fine level)
    OP_UPDATE,       // u64 += (double) u[src] over the shape (level 0)
    OP_RESIDUAL64,   // level-0 f32 = (float)(f64 - L u64) on the interior, max |r| per problem
    OP_FNORM,        // max |f64| per problem over the shape
};

struct BArgs {
    const BGrid* grids;
    const BTop* tops;
    BState* state;
    const BItem* items;
    int32_t nitems;
    int32_t op;
    int32_t src, src2;  // current-buffer index (of the level, and of the coarse level)
    int32_t colour;
};

__device__ __forceinline__ double wave_max64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8), each
// with its own L2.  Neighbouring items of a list share halo rows, so each XCD
// takes one contiguous eighth of the list: its workgroups walk that range
// with a stride of their own count, and the items in flight on one XCD are
// neighbours.  Grids not a multiple of 8 walk the whole list round-robin.
struct Walk {
    int first, end, step;
};
__device__ __forceinline__ Walk xcd_walk(int nitems) {
    const int G = (int)gridDim.x, b = (int)blockIdx.x;
    if (G % 8) return Walk{b, nitems, G};
    const int per = G / 8, x = b % 8;
    const int lo = (int)((int64_t)nitems * x / 8), hi = (int)((int64_t)nitems * (x + 1) / 8);
    return Walk{lo + b / 8, hi, per};
}
// this workgroup's place when workgroups are numbered XCD by XCD
__device__ __forceinline__ int xcd_rank() {
    const int G = (int)gridDim.x, b = (int)blockIdx.x;
    return G % 8 ? b : (b % 8) * (G / 8) + b / 8;
}

// max |.| of the FP64 residual or of f per problem: each workgroup walks a
// contiguous run of items (the lists hold each problem's items together) and
// keeps the running max of the problem at hand, folding it into the
// problem's norm (one atomic per wave) only when the problem changes, instead
// of once per item
template <int OP>
__device__ __forceinline__ void batch_norms(const BArgs& a) {
    const int per = (a.nitems + (int)gridDim.x - 1) / (int)gridDim.x;
    const int i0 = xcd_rank() * per, i1 = min(a.nitems, i0 + per);
    int cur = -1;
    double vmax = 0.0;
    auto flush = [&]() {
        vmax = wave_max64(vmax);
        if ((threadIdx.x & 63) == 0 && vmax > 0.0) {
            unsigned long long* dst = OP == OP_FNORM ? (unsigned long long*)&a.state[cur].fnorm : &a.state[cur].rnorm;
            atomicMax(dst, (unsigned long long)__double_as_longlong(vmax));
        }
        vmax = 0.0;
    };
    for (int it = i0; it < i1; it++) {
        const BItem item = a.items[it];
        const int p = a.grids[item.grid].problem;
        if (p != cur) {  // uniform per workgroup
            if (cur >= 0) flush();
            cur = p;
        }
        if (!a.state[p].active) continue;
        const BTop& t = a.tops[p];
        const int64_t k = item.z0 + threadIdx.x;
        const int64_t kz1 = t.u.n2;
        if (OP == OP_RESIDUAL64 ? (k < kz1 - 1 && k >= 1) : k < kz1) {
            for (int64_t j = item.y0 + threadIdx.y; j < item.y1; j += blockDim.y) {
                for (int64_t i = item.x0; i < item.x1; i++) {
                    if (OP == OP_RESIDUAL64) {
                        const double rv = residual_point<double, MGP_MODE_CONSISTENT>(t.u, t.u.ix(i, j, k),
                                                                                      t.f.at(i, j, k), t.hsq, t.inv_h2);
                        a.grids[item.grid].f.at(i, j, k) = (float)rv;  // the next iteration's right-hand side
                        vmax = fmax(vmax, fabs(rv));
                    } else {
                        vmax = fmax(vmax, fabs(t.f.at(i, j, k)));
                    }
                }
            }
        }
    }
    if (cur >= 0) flush();
}

// Jacobi sweep or residual of one item's column k: march along x holding
// u(i-1), u(i), u(i+1) in registers, so each point loads one new x-plane value
// instead of three (and the restrict-qualified pointers let the next point's
// loads issue before this point's store)
template <int OP>
__device__ __forceinline__ void march_item(const BGrid& g, const BItem& item, int src, int64_t k) {
    const View<float>& v = g.u[src];
    const View<float>& w = OP == OP_JACOBI ? g.u[1 - src] : g.r;
    const float* __restrict__ q = v.p;
    const float* __restrict__ fp = g.f.p;
    float* __restrict__ o = w.p;
    const int64_t s0 = v.s0, s1 = v.s1;
    for (int64_t j = item.y0 + threadIdx.y; j < item.y1; j += blockDim.y) {
        int64_t c = v.ix(item.x0, j, k), cf = g.f.ix(item.x0, j, k), co = w.ix(item.x0, j, k);
        float um = q[c - s0], uc = q[c];
        for (int64_t i = item.x0; i < item.x1; i++) {
            const float up = q[c + s0];
            const float uS = q[c - s1], uN = q[c + s1], uB = q[c - 1], uT = q[c + 1], fc = fp[cf];
            if (OP == OP_JACOBI) {
                o[co] = relax_point<float, MGP_MODE_CONSISTENT>(uc, um, up, uS, uN, uB, uT, fc, g.one_omega,
                                                                g.omega_sixth, g.hsq);
            } else {  // residual_point's consistent form
                const float sum = ((um - uc) + (up - uc)) + ((uS - uc) + (uN - uc)) + ((uB - uc) + (uT - uc));
                o[co] = fma(-g.inv_h2, sum, fc);
            }
            um = uc;
            uc = up;
            c += s0;
            cf += g.f.s0;
            co += w.s0;
        }
    }
}

template <int OP>
__global__ void __launch_bounds__(256) k_batch(BArgs a) {
    if (OP == OP_RESIDUAL64 || OP == OP_FNORM) {
        batch_norms<OP>(a);
        return;
    }
    const Walk w = xcd_walk(a.nitems);
    if (OP == OP_JACOBI || OP == OP_RESIDUAL) {
        for (int it = w.first; it < w.end; it += w.step) {
            const BItem item = a.items[it];
            const BGrid& g = a.grids[item.grid];
            if (!a.state[g.problem].active) continue;  // uniform per workgroup
            const int64_t k = item.z0 + threadIdx.x;
            if (k >= 1 && k < g.u[a.src].n2 - 1) march_item<OP>(g, item, a.src, k);
        }
        return;
    }
    for (int it = w.first; it < w.end; it += w.step) {
        const BItem item = a.items[it];
        const BGrid& g = a.grids[item.grid];
        const int p = g.problem;
        if (!a.state[p].active) continue;  // uniform per workgroup
        const int64_t k = item.z0 + threadIdx.x;
        const View<float>& us = g.u[a.src];
        const int64_t kz1 = us.n2;  // level 0 of the FP64 arrays has the level-0 shape
        // interior-only operations stop one short of the far face
        constexpr bool interior = OP == OP_GS_HALF;  // Jacobi and the residual: march_item
        const int64_t kmax = interior ? kz1 - 1 : kz1;
        if (k < kmax && (!interior || k >= 1)) {
            for (int64_t j = item.y0 + threadIdx.y; j < item.y1; j += blockDim.y) {
                for (int64_t i = item.x0; i < item.x1; i++) {
                    switch (OP) {
                    case OP_GS_HALF: {
                        if ((int)((i + j + k) & 1) != a.colour) break;
                        const int64_t c = us.ix(i, j, k);
                        float* q = us.p;
                        q[c] = relax_point<float, MGP_MODE_CONSISTENT>(q[c], q[c - us.s0], q[c + us.s0],
                                                                       q[c - us.s1], q[c + us.s1], q[c - 1], q[c + 1],
                                                                       g.f.at(i, j, k), g.one_omega, g.omega_sixth,
                                                                       g.hsq);
                        break;
                    }
                    case OP_FW: {
                        const BGrid& fg = a.grids[g.fine];
                        g.f.at(i, j, k) = fw_point<float>(fg.r, i, j, k);
                        break;
                    }
                    case OP_ZERO_U:
                        us.at(i, j, k) = 0.0f;
                        break;
                    case OP_PROLONG_ADD: {
                        // grid = fine level; its coarse level is the grid that names it as `fine`
                        const BGrid& cg = a.grids[item.grid + 1];
                        us.at(i, j, k) += trilinear<float>(cg.u[a.src2], i, j, k);
                        break;
                    }
                    case OP_UPDATE:
                        a.tops[p].u.at(i, j, k) += (double)us.at(i, j, k);
                        break;
                    default:  // the norms: batch_norms
                        break;
                    }
                }
            }
        }
    }
}

// Refinement stop rule per problem, the one RefinementSolver::solve applies:
// stop at norm <= tol; after miniter iterations also when the norm more than
// doubles; at maxiter iterations.  `first` records the initial norm.
__global__ void k_batch_check(BState* st, int n, double tol, int miniter, int maxiter, int first, int* remaining) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !st[p].active) return;
    const double norm = __longlong_as_double((long long)st[p].rnorm) / st[p].fnorm;
    st[p].rnorm = 0ull;
    if (first) {
        st[p].rel = norm;
        if (maxiter <= 0) st[p].active = 0;
    } else if (norm <= tol) {
        st[p].rel = norm;
        st[p].active = 0;
    } else if (st[p].iter >= miniter && norm > 2.0 * st[p].rel) {
        st[p].rel = norm;
        st[p].active = 0;
    } else {
        st[p].rel = norm;
        st[p].iter += 1;
        if (st[p].iter >= maxiter) st[p].active = 0;
    }
    if (st[p].active) atomicAdd(remaining, 1);
}

__global__ void k_batch_reset(BState* st, int n, int* remaining) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) *remaining = 0;
    if (p >= n) return;
    st[p].rnorm = 0ull;
    st[p].fnorm = 0.0;
    st[p].rel = 0.0;
    st[p].iter = 0;
    st[p].active = 1;
}

__global__ void k_zero_int(int* p) { *p = 0; }

// -------------------------------------------------------------- host side

struct Tile {
    int xd, yd;  // planes per item, rows per item
};

// workgroups of a batched launch (fewer when the list has fewer items): the
// items are small, so many workgroups keep enough loads in flight
int64_t g_batch_blocks = 16384;

static std::mutex g_tile_mu;
static std::map<std::tuple<int64_t, int64_t, int64_t>, Tile> g_tile_cache;

static int check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

class BatchSolver {
public:
    BatchSolver(const mgp_batch_config& c, const int64_t* dims) : cfg(c) {
        if (c.count < 1 || c.nsmooth < 1 || c.maxiter < 0 || c.miniter < 0 || !(c.tolerance >= 0.0) ||
            (c.relaxation != MGP_JACOBI && c.relaxation != MGP_GAUSS_SEIDEL) || c.coarse_sweeps < 0)
            throw std::invalid_argument("invalid batch config");
        omega = c.omega > 0.0 ? c.omega : (c.relaxation == MGP_JACOBI ? 6.0 / 7.0 : 8.0 / 9.0);
        n = (int)c.count;
        shapes.assign(dims, dims + 3 * n);
        plan();
        allocate();
        build_lists();
        check_hip(hipStreamCreate(&stream), "hipStreamCreate");
    }
    ~BatchSolver() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second.exec);
        for (BItem* d : lists_owned)
            if (d) (void)hipFree(d);
        if (stream) (void)hipStreamSynchronize(stream), (void)hipStreamDestroy(stream);
        if (arena) (void)hipFree(arena);
        if (host_remaining) (void)hipHostFree(host_remaining);
    }

    void set_rhs(int64_t p, const double* f) {
        check_p(p);
        const BTop& t = htops[p];
        const int64_t cnt = count_of(p, 0);
        check_hip(hipMemcpyAsync(t.f.p, f, cnt * sizeof(double), hipMemcpyHostToDevice, stream), "set_rhs");
        check_hip(hipStreamSynchronize(stream), "set_rhs");
    }
    void get_solution(int64_t p, double* out) {
        check_p(p);
        check_hip(hipMemcpyAsync(out, htops[p].u.p, count_of(p, 0) * sizeof(double), hipMemcpyDeviceToHost, stream),
                  "get_solution");
        check_hip(hipStreamSynchronize(stream), "get_solution");
    }
    void stats(int64_t p, double* rel, int64_t* iters) {
        check_p(p);
        BState s;
        check_hip(hipMemcpy(&s, state + p, sizeof(BState), hipMemcpyDeviceToHost), "stats");
        if (rel) *rel = s.rel;
        if (iters) *iters = s.iter;
    }
    void tile(int64_t p, int64_t level, int64_t out[2]) {
        check_p(p);
        if (level < 0 || level >= levels[p]) throw std::invalid_argument("level out of range");
        const Tile t = tiles[grid_index(p, level)];
        out[0] = t.xd;
        out[1] = t.yd;
    }

    // u = 0 on every level, then refinement iterations to the tolerance
    double solve() {
        const auto t0 = std::chrono::steady_clock::now();
        // u64 = 0 (zero Dirichlet data), state reset, ||f||, initial residual
        check_hip(hipMemsetAsync(top_u_base, 0, top_u_bytes, stream), "solve");
        hipLaunchKernelGGL(k_batch_reset, dim3((n + 255) / 256), dim3(256), 0, stream, state, n, dev_remaining);
        for (auto& c : cur) std::fill(c.begin(), c.end(), 0);
        launch(OP_FNORM, list_top, 0, 0, 0);
        launch(OP_RESIDUAL64, list_top_interior, 0, 0, 0);
        hipLaunchKernelGGL(k_batch_check, dim3((n + 255) / 256), dim3(256), 0, stream, state, n, cfg.tolerance,
                           (int)cfg.miniter, (int)cfg.maxiter, 1, dev_remaining);
        for (int64_t it = 0; it < cfg.maxiter; it++) {
            run_iteration();
            check_hip(hipMemcpyAsync(host_remaining, dev_remaining, sizeof(int), hipMemcpyDeviceToHost, stream),
                      "solve");
            check_hip(hipStreamSynchronize(stream), "solve");
            if (*host_remaining == 0) break;
        }
        check_hip(hipStreamSynchronize(stream), "solve");
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

private:
    mgp_batch_config cfg;
    double omega;
    int n;
    std::vector<int64_t> shapes;  // n x 3
    std::vector<int> levels;      // per problem
    std::vector<int> grid0;       // first grid index of each problem
    std::vector<std::array<int64_t, 3>> gshape;  // per grid
    std::vector<BGrid> hgrids;
    std::vector<BTop> htops;
    std::vector<Tile> tiles;
    int maxlev = 0;
    char* arena = nullptr;
    void* top_u_base = nullptr;
    size_t top_u_bytes = 0;
    BGrid* dgrids = nullptr;
    BTop* dtops = nullptr;
    BState* state = nullptr;
    int* dev_remaining = nullptr;
    int* host_remaining = nullptr;
    hipStream_t stream = nullptr;
    // work lists: per level, interior tiles / shape tiles of the problems that
    // go below that level (A) and of the problems whose coarsest level it is (B)
    struct List {
        BItem* dev = nullptr;
        int count = 0;
    };
    std::vector<List> relax_a, relax_b, shape_a, shape_b, fw_to;  // fw_to[l]: coarse tiles of level l (group A of l-1)
    List list_top, list_top_interior, list_l0;
    std::vector<std::vector<int>> cur;  // [level][0: group A, 1: group B] current buffer
    struct Replay {
        hipGraphExec_t exec;
        std::vector<std::vector<int>> after;  // buffer state after the recorded iteration
    };
    std::map<std::vector<int>, Replay> graphs;

    void check_p(int64_t p) const {
        if (p < 0 || p >= n) throw std::invalid_argument("problem index out of range");
    }
    int grid_index(int64_t p, int64_t l) const { return grid0[p] + (int)l; }
    int64_t count_of(int64_t p, int l) const {
        const auto& s = gshape[grid_index(p, l)];
        return s[0] * s[1] * s[2];
    }

    void plan() {
        levels.resize(n);
        grid0.resize(n);
        int g = 0;
        for (int p = 0; p < n; p++) {
            const int64_t* d = &shapes[3 * p];
            int64_t m = std::min({d[0], d[1], d[2]});
            if (m < 3) throw std::invalid_argument("every axis needs at least 3 points");
            int L = 1;
            // coarsen while every axis stays 2^k+1-compatible and above 3 points
            while (true) {
                bool ok = true;
                for (int a = 0; a < 3; a++) {
                    const int64_t s = (d[a] - 1) >> (L - 1);
                    if (s % 2 != 0 || s / 2 < 2) ok = false;
                }
                if (!ok) break;
                L++;
            }
            if (cfg.levels > 0) {
                if (cfg.levels > L) throw std::invalid_argument("too many levels for a problem's shape");
                L = (int)cfg.levels;
            }
            levels[p] = L;
            grid0[p] = g;
            for (int l = 0; l < L; l++)
                gshape.push_back({((d[0] - 1) >> l) + 1, ((d[1] - 1) >> l) + 1, ((d[2] - 1) >> l) + 1});
            g += L;
            maxlev = std::max(maxlev, L);
        }
        cur.assign(maxlev, std::vector<int>(2, 0));
    }

    static size_t align(size_t x) { return (x + 255) & ~(size_t)255; }

    void allocate() {
        const int ng = (int)gshape.size();
        size_t bytes = 0;
        std::vector<size_t> off32(ng), off64(n);
        // FP64 u of every problem first (one contiguous block: zeroed per solve)
        for (int p = 0; p < n; p++) {
            off64[p] = bytes;
            bytes += align(count_of(p, 0) * sizeof(double));
        }
        top_u_bytes = bytes;
        const size_t fr64 = bytes;
        for (int p = 0; p < n; p++) bytes += align(count_of(p, 0) * sizeof(double));
        for (int g = 0; g < ng; g++) {
            off32[g] = bytes;
            const int64_t c = gshape[g][0] * gshape[g][1] * gshape[g][2];
            bytes += 4 * align(c * sizeof(float));
        }
        const size_t tab = bytes;
        bytes += align(ng * sizeof(BGrid)) + align(n * sizeof(BTop)) + align(n * sizeof(BState)) + 256;
        check_hip(hipMalloc((void**)&arena, bytes), "batch arena");
        check_hip(hipMemset(arena, 0, bytes), "batch arena");
        top_u_base = arena;
        auto view64 = [&](size_t off, const std::array<int64_t, 3>& s) {
            View<double> v;
            v.p = (double*)(arena + off);
            v.n0 = s[0], v.n1 = s[1], v.n2 = s[2], v.s0 = s[1] * s[2], v.s1 = s[2];
            return v;
        };
        auto view32 = [&](size_t off, const std::array<int64_t, 3>& s) {
            View<float> v;
            v.p = (float*)(arena + off);
            v.n0 = s[0], v.n1 = s[1], v.n2 = s[2], v.s0 = s[1] * s[2], v.s1 = s[2];
            return v;
        };
        htops.resize(n);
        size_t o = fr64;
        for (int p = 0; p < n; p++) {
            const auto& s = gshape[grid0[p]];
            const size_t sz = align(count_of(p, 0) * sizeof(double));
            BTop t;
            t.u = view64(off64[p], s);
            t.f = view64(o, s);
            o += sz;
            const double h = 1.0 / (double)(shapes[3 * p] - 1);
            t.hsq = h * h;  // include/residual.h:23
            t.inv_h2 = 1.0 / (h * h);
            htops[p] = t;
        }
        hgrids.resize(ng);
        for (int p = 0; p < n; p++) {
            const double h0 = 1.0 / (double)(shapes[3 * p] - 1);
            for (int l = 0; l < levels[p]; l++) {
                const int g = grid_index(p, l);
                const auto& s = gshape[g];
                const size_t sz = align(s[0] * s[1] * s[2] * sizeof(float));
                BGrid b;
                b.u[0] = view32(off32[g], s);
                b.u[1] = view32(off32[g] + sz, s);
                b.f = view32(off32[g] + 2 * sz, s);
                b.r = view32(off32[g] + 3 * sz, s);
                b.problem = p;
                b.fine = l > 0 ? g - 1 : -1;
                const double h = h0 * (double)(1 << l);  // level spacing h * 2^l
                const RelaxConsts<float> rc(omega, h);
                b.one_omega = rc.one_omega;
                b.omega_sixth = rc.omega_sixth;
                b.hsq = rc.hsq;
                b.inv_h2 = (float)(1.0 / (h * h));
                hgrids[g] = b;
            }
        }
        dgrids = (BGrid*)(arena + tab);
        dtops = (BTop*)(arena + tab + align(ng * sizeof(BGrid)));
        state = (BState*)(arena + tab + align(ng * sizeof(BGrid)) + align(n * sizeof(BTop)));
        dev_remaining = (int*)(arena + tab + align(ng * sizeof(BGrid)) + align(n * sizeof(BTop)) +
                               align(n * sizeof(BState)));
        check_hip(hipMemcpy(dgrids, hgrids.data(), ng * sizeof(BGrid), hipMemcpyHostToDevice), "batch tables");
        check_hip(hipMemcpy(dtops, htops.data(), n * sizeof(BTop), hipMemcpyHostToDevice), "batch tables");
        check_hip(hipHostMalloc((void**)&host_remaining, sizeof(int)), "batch");
    }

    // tiles of one grid's region [lo, hi) per axis with the grid's tile shape
    static void tiles_of(std::vector<BItem>& out, int g, const int64_t lo[3], const int64_t hi[3], Tile t) {
        for (int64_t x = lo[0]; x < hi[0]; x += t.xd)
            for (int64_t y = lo[1]; y < hi[1]; y += t.yd)
                for (int64_t z = lo[2]; z < hi[2]; z += 64) {
                    BItem it;
                    it.grid = g;
                    it.x0 = (int32_t)x;
                    it.x1 = (int32_t)std::min(hi[0], x + t.xd);
                    it.y0 = (int32_t)y;
                    it.y1 = (int32_t)std::min(hi[1], y + t.yd);
                    it.z0 = (int32_t)z;
                    out.push_back(it);
                }
    }
    std::vector<BItem> region(int g, bool interior, Tile t) const {
        const auto& s = gshape[g];
        const int64_t lo[3] = {interior ? 1 : 0, interior ? 1 : 0, 0};  // z bounds are applied per thread
        const int64_t hi[3] = {interior ? s[0] - 1 : s[0], interior ? s[1] - 1 : s[1], s[2]};
        std::vector<BItem> v;
        tiles_of(v, g, lo, hi, t);
        return v;
    }
    List upload(const std::vector<BItem>& v) {
        List L;
        L.count = (int)v.size();
        if (L.count) {
            check_hip(hipMalloc((void**)&L.dev, v.size() * sizeof(BItem)), "work list");
            check_hip(hipMemcpy(L.dev, v.data(), v.size() * sizeof(BItem), hipMemcpyHostToDevice), "work list");
        }
        lists_owned.push_back(L.dev);
        return L;
    }
    std::vector<BItem*> lists_owned;

    void autotune() {
        // candidate tiles, timed per distinct level shape on a Jacobi sweep of
        // every grid of that shape
        static const Tile cands[] = {{1, 4},  {2, 4}, {4, 4}, {8, 4},  {16, 4}, {1, 8},
                                     {2, 8}, {4, 8}, {8, 8}, {1, 16}, {4, 16}};
        const int ng = (int)gshape.size();
        tiles.assign(ng, Tile{1, 4});
        std::map<std::tuple<int64_t, int64_t, int64_t>, std::vector<int>> by_shape;
        for (int g = 0; g < ng; g++) by_shape[{gshape[g][0], gshape[g][1], gshape[g][2]}].push_back(g);
        hipStream_t s;
        check_hip(hipStreamCreate(&s), "autotune");
        hipEvent_t e0, e1;
        check_hip(hipEventCreate(&e0), "autotune");
        check_hip(hipEventCreate(&e1), "autotune");
        // every problem active while timing
        std::vector<BState> on(n);
        for (auto& x : on) x.active = 1, x.rnorm = 0, x.fnorm = 1.0, x.rel = 0.0, x.iter = 0;
        check_hip(hipMemcpy(state, on.data(), n * sizeof(BState), hipMemcpyHostToDevice), "autotune");
        for (auto& kv : by_shape) {
            Tile best{1, 4};
            bool cached = false;
            {
                std::lock_guard<std::mutex> lk(g_tile_mu);
                auto it = g_tile_cache.find(kv.first);
                if (it != g_tile_cache.end() && !(cfg.flags & MGP_BATCH_RETUNE)) best = it->second, cached = true;
            }
            if (!cached) {
                float best_ms = 1e30f;
                for (const Tile& t : cands) {
                    std::vector<BItem> v;
                    for (int g : kv.second) {
                        auto r = region(g, true, t);
                        v.insert(v.end(), r.begin(), r.end());
                    }
                    if (v.empty()) continue;
                    BItem* d = nullptr;
                    check_hip(hipMalloc((void**)&d, v.size() * sizeof(BItem)), "autotune");
                    check_hip(hipMemcpy(d, v.data(), v.size() * sizeof(BItem), hipMemcpyHostToDevice), "autotune");
                    BArgs a = args(OP_JACOBI, d, (int)v.size(), 0, 0, 0);
                    const dim3 grid(persistent_blocks((int)v.size()));
                    launch_op(OP_JACOBI, grid, s, a);  // warm
                    check_hip(hipEventRecord(e0, s), "autotune");
                    for (int r = 0; r < 3; r++) launch_op(OP_JACOBI, grid, s, a);
                    check_hip(hipEventRecord(e1, s), "autotune");
                    check_hip(hipEventSynchronize(e1), "autotune");
                    float ms = 0.f;
                    check_hip(hipEventElapsedTime(&ms, e0, e1), "autotune");
                    (void)hipFree(d);
                    if (ms < best_ms) best_ms = ms, best = t;
                }
                std::lock_guard<std::mutex> lk(g_tile_mu);
                g_tile_cache[kv.first] = best;
            }
            for (int g : kv.second) tiles[g] = best;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        // the timing sweeps wrote into the u buffers: clear them again
        for (const BGrid& b : hgrids) {
            const size_t c = b.u[0].n0 * b.u[0].n1 * b.u[0].n2 * sizeof(float);
            check_hip(hipMemset(b.u[0].p, 0, c), "autotune");
            check_hip(hipMemset(b.u[1].p, 0, c), "autotune");
        }
    }

    void build_lists() {
        autotune();
        relax_a.resize(maxlev);
        relax_b.resize(maxlev);
        shape_a.resize(maxlev);
        shape_b.resize(maxlev);
        fw_to.resize(maxlev);
        for (int l = 0; l < maxlev; l++) {
            std::vector<BItem> ra, rb, sa, sb, fw;
            for (int p = 0; p < n; p++) {
                if (l >= levels[p]) continue;
                const int g = grid_index(p, l);
                const Tile t = tiles[g];
                auto ri = region(g, true, t);
                auto rs = region(g, false, t);
                if (l < levels[p] - 1) {
                    ra.insert(ra.end(), ri.begin(), ri.end());
                    sa.insert(sa.end(), rs.begin(), rs.end());
                } else {
                    rb.insert(rb.end(), ri.begin(), ri.end());
                    sb.insert(sb.end(), rs.begin(), rs.end());
                }
                if (l > 0) fw.insert(fw.end(), rs.begin(), rs.end());  // coarse tiles of level l
            }
            relax_a[l] = upload(ra);
            relax_b[l] = upload(rb);
            shape_a[l] = upload(sa);
            shape_b[l] = upload(sb);
            fw_to[l] = upload(fw);
        }
        std::vector<BItem> top, topi, l0;
        for (int p = 0; p < n; p++) {
            const int g = grid_index(p, 0);
            auto rs = region(g, false, tiles[g]);
            auto ri = region(g, true, tiles[g]);
            top.insert(top.end(), rs.begin(), rs.end());
            topi.insert(topi.end(), ri.begin(), ri.end());
        }
        list_top = upload(top);
        list_top_interior = upload(topi);
    }

    static int persistent_blocks(int nitems) {
        const int64_t g = std::max<int64_t>(1, std::min<int64_t>(nitems, g_batch_blocks));
        return (int)(g >= 64 ? (g + 7) / 8 * 8 : g);  // a multiple of 8 walks XCD by XCD
    }

    BArgs args(int op, const BItem* items, int nitems, int src, int src2, int colour) const {
        BArgs a;
        a.grids = dgrids;
        a.tops = dtops;
        a.state = state;
        a.items = items;
        a.nitems = nitems;
        a.op = op;
        a.src = src;
        a.src2 = src2;
        a.colour = colour;
        return a;
    }
    static void launch_op(int op, dim3 grid, hipStream_t s, const BArgs& a) {
        switch (op) {
        case OP_JACOBI: hipLaunchKernelGGL(k_batch<OP_JACOBI>, grid, dim3(64, 4), 0, s, a); break;
        case OP_GS_HALF: hipLaunchKernelGGL(k_batch<OP_GS_HALF>, grid, dim3(64, 4), 0, s, a); break;
        case OP_RESIDUAL: hipLaunchKernelGGL(k_batch<OP_RESIDUAL>, grid, dim3(64, 4), 0, s, a); break;
        case OP_FW: hipLaunchKernelGGL(k_batch<OP_FW>, grid, dim3(64, 4), 0, s, a); break;
        case OP_ZERO_U: hipLaunchKernelGGL(k_batch<OP_ZERO_U>, grid, dim3(64, 4), 0, s, a); break;
        case OP_PROLONG_ADD: hipLaunchKernelGGL(k_batch<OP_PROLONG_ADD>, grid, dim3(64, 4), 0, s, a); break;
        case OP_UPDATE: hipLaunchKernelGGL(k_batch<OP_UPDATE>, grid, dim3(64, 4), 0, s, a); break;
        case OP_RESIDUAL64: hipLaunchKernelGGL(k_batch<OP_RESIDUAL64>, grid, dim3(64, 4), 0, s, a); break;
        default: hipLaunchKernelGGL(k_batch<OP_FNORM>, grid, dim3(64, 4), 0, s, a); break;
        }
    }
    void launch(int op, const List& L, int src, int src2, int colour) {
        if (!L.count) return;
        launch_op(op, dim3(persistent_blocks(L.count)), stream, args(op, L.dev, L.count, src, src2, colour));
        check_hip(hipGetLastError(), "batch launch");
    }

    // nsweeps smoothing passes over list L at level l, group grp
    void relax(const List& L, int l, int grp, int64_t nsweeps) {
        for (int64_t s = 0; s < nsweeps; s++) {
            if (cfg.relaxation == MGP_JACOBI) {
                launch(OP_JACOBI, L, cur[l][grp], 0, 0);
                cur[l][grp] ^= 1;
            } else {
                launch(OP_GS_HALF, L, cur[l][grp], 0, 0);
                launch(OP_GS_HALF, L, cur[l][grp], 0, 1);
            }
        }
    }

    // one refinement iteration: e = 0, one FP32 V-cycle on f32 (the previous
    // residual, already FP32), u64 += e, the residual into f32 and its norm,
    // the stop rule
    void enqueue_iteration() {
        const int64_t nu = cfg.nsmooth;
        const int64_t cs = cfg.coarse_sweeps > 0 ? cfg.coarse_sweeps : nu;
        // level-0 e lives in the current buffer of level 0 (group A, or B for
        // one-level problems)
        launch(OP_ZERO_U, shape_a[0], cur[0][0], 0, 0);
        launch(OP_ZERO_U, shape_b[0], cur[0][1], 0, 0);
        for (int l = 0; l < maxlev; l++) {
            relax(relax_a[l], l, 0, nu);
            launch(OP_RESIDUAL, relax_a[l], cur[l][0], 0, 0);
            if (l + 1 < maxlev) {
                launch(OP_FW, fw_to[l + 1], 0, 0, 0);
                // coarse u = 0: the group a problem belongs to on level l+1
                launch(OP_ZERO_U, shape_a[l + 1], cur[l + 1][0], 0, 0);
                launch(OP_ZERO_U, shape_b[l + 1], cur[l + 1][1], 0, 0);
            }
            relax(relax_b[l], l, 1, cs);
        }
        for (int l = maxlev - 2; l >= 0; l--) {
            // coarse level l+1: group A or B, per problem; the prolongation
            // reads the coarse buffer of the problem's group there
            prolong_add(l);
            relax(relax_a[l], l, 0, nu);
        }
        update_level0();
        launch(OP_RESIDUAL64, list_top_interior, 0, 0, 0);
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, stream, dev_remaining);
        hipLaunchKernelGGL(k_batch_check, dim3((n + 255) / 256), dim3(256), 0, stream, state, n, cfg.tolerance,
                           (int)cfg.miniter, (int)cfg.maxiter, 0, dev_remaining);
    }
    void prolong_add(int l) {
        // fine tiles of problems in group A at level l, split by the coarse
        // group (A or B on level l+1) because the coarse buffer index differs
        if (!prolong_lists_built) build_prolong_lists();
        launch(OP_PROLONG_ADD, prolong_aa[l], cur[l][0], cur[l + 1][0], 0);
        launch(OP_PROLONG_ADD, prolong_ab[l], cur[l][0], cur[l + 1][1], 0);
    }
    void update_level0() {
        if (!prolong_lists_built) build_prolong_lists();
        launch(OP_UPDATE, top_a, cur[0][0], 0, 0);
        launch(OP_UPDATE, top_b, cur[0][1], 0, 0);
    }
    bool prolong_lists_built = false;
    std::vector<List> prolong_aa, prolong_ab;
    List top_a, top_b;
    void build_prolong_lists() {
        prolong_aa.resize(maxlev);
        prolong_ab.resize(maxlev);
        for (int l = 0; l + 1 < maxlev; l++) {
            std::vector<BItem> aa, ab;
            for (int p = 0; p < n; p++) {
                if (l + 1 >= levels[p]) continue;
                const int g = grid_index(p, l);
                auto rs = region(g, false, tiles[g]);
                auto& dst = (l + 1 < levels[p] - 1) ? aa : ab;
                dst.insert(dst.end(), rs.begin(), rs.end());
            }
            prolong_aa[l] = upload(aa);
            prolong_ab[l] = upload(ab);
        }
        std::vector<BItem> ta, tb;
        for (int p = 0; p < n; p++) {
            const int g = grid_index(p, 0);
            auto rs = region(g, false, tiles[g]);
            auto& dst = levels[p] > 1 ? ta : tb;
            dst.insert(dst.end(), rs.begin(), rs.end());
        }
        top_a = upload(ta);
        top_b = upload(tb);
        prolong_lists_built = true;
    }

    void run_iteration() {
        if (!prolong_lists_built) build_prolong_lists();
        std::vector<int> key;
        for (auto& c : cur) key.push_back(c[0]), key.push_back(c[1]);
        if (cfg.flags & MGP_BATCH_NO_GRAPHS) {
            enqueue_iteration();
            return;
        }
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            const auto before = cur;
            hipGraph_t g;
            check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "capture");
            enqueue_iteration();
            check_hip(hipStreamEndCapture(stream, &g), "capture");
            hipGraphExec_t ex;
            check_hip(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "instantiate");
            (void)hipGraphDestroy(g);
            it = graphs.emplace(key, Replay{ex, cur}).first;
            cur = before;
        }
        check_hip(hipGraphLaunch(it->second.exec, stream), "graph launch");
        cur = it->second.after;
    }

};

}  // namespace mgp

// ============================================================ C entry points

using mgp::BatchSolver;

namespace mgp {
void set_last_error(const char* msg);  // mgp_solver.cpp: the text mgp_last_error returns
}

template <class F>
static int batch_guarded(F&& f) {
    try {
        f();
        return MGP_OK;
    } catch (const std::invalid_argument& e) {
        mgp::set_last_error(e.what());
        return MGP_ERR_ARG;
    } catch (const std::exception& e) {
        mgp::set_last_error(e.what());
        return MGP_ERR_STATE;
    }
}

extern "C" {

int mgp_batch_create(const mgp_batch_config* c, const int64_t* dims, void** out) {
    if (!c || !dims || !out) return MGP_ERR_ARG;
    return batch_guarded([&] { *out = new BatchSolver(*c, dims); });
}
int mgp_batch_destroy(void* h) {
    delete static_cast<BatchSolver*>(h);
    return MGP_OK;
}
int mgp_batch_set_rhs(void* h, int64_t p, const double* f) {
    if (!h || !f) return MGP_ERR_ARG;
    return batch_guarded([&] { static_cast<BatchSolver*>(h)->set_rhs(p, f); });
}
int mgp_batch_solve(void* h, double* seconds) {
    if (!h) return MGP_ERR_ARG;
    return batch_guarded([&] {
        const double t = static_cast<BatchSolver*>(h)->solve();
        if (seconds) *seconds = t;
    });
}
int mgp_batch_stats(void* h, int64_t p, double* rel_res, int64_t* iterations) {
    if (!h) return MGP_ERR_ARG;
    return batch_guarded([&] { static_cast<BatchSolver*>(h)->stats(p, rel_res, iterations); });
}
int mgp_batch_solution(void* h, int64_t p, double* out) {
    if (!h || !out) return MGP_ERR_ARG;
    return batch_guarded([&] { static_cast<BatchSolver*>(h)->get_solution(p, out); });
}
int mgp_batch_tile(void* h, int64_t p, int64_t level, int64_t out[2]) {
    if (!h || !out) return MGP_ERR_ARG;
    return batch_guarded([&] { static_cast<BatchSolver*>(h)->tile(p, level, out); });
}

}  // extern "C"
