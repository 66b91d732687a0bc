// One-pass red-black Gauss-Seidel sweep for gfx950 (mgp_rbgs_sweep_*).
//
// The reference runs a sweep as two in-place half-sweeps over the whole
// level, each preceded by a Neumann ghost refresh (include/relaxation.h:25-90,
// include/gaussseidel.h:28-94): every point is read and written twice.  Here a
// workgroup owns a (y, z) tile and marches along x; at plane i it updates the
// red points of plane i (on the tile plus a one-point ring, from the old
// values) and then the black points of plane i-1 on the tile (from the new
// red values), writing both colours of plane i-1 to the output buffer.  The
// ring's red values are recomputed by the neighbouring tiles as well, so tiles
// never wait for each other; the output goes to a second buffer so no tile
// can see another tile's new values early.  Per point the arithmetic is the
// reference's (or the consistent mode's) update, and ghost values are formed
// exactly as the ghost refresh would have stored them, so the result equals
// the two half-sweeps bit for bit.
//
// Tile: 64 z-columns by 16 y-rows of ring positions per plane (62 x 14
// interior); 512 threads (8 waves), each owning 2 ring positions (2 rows of one
// column) in registers.  LDS holds the old plane (with a 2-point ring, for the
// red stencil) and two planes of new red values (ping-pong).  The loop body is
// branch-free per lane: positions outside the box keep their value through a
// select, Neumann ghosts are folded into the loaded values (old plane) or
// written into LDS by the ghost positions' owners (new red plane), and every
// load is unconditional from a valid address, so the compiler can count
// outstanding memory operations exactly and the next plane's loads overlap a
// whole step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "mgp.h"
#include "mgp_device.h"
#include "mgp_point.h"
#include "mgp_sweep.h"

namespace mgp {

// waves per launch the residual, fused residual + restriction and
// prolongation marches aim for (0: "sweep.target_wgs"; the residual's only
// where it stores nothing but its norm).  Measured at
// 1025 x 1025 x 513 (FP64): the fused residual + restriction 2.81 ms with
// 16384 waves, 2.60 ms with 32768 (no gain beyond); the prolongation 2.26 ms
// with 16384, 2.04 ms with 65536; the norm-only residual 1.59 ms with 16384,
// 1.51 ms with 4096 (the stored residual 3.31 ms with 16384, 3.49 ms with
// 4096).  Coarser levels are flat in all three.
int64_t g_resid_target = 4096;
int64_t g_fw_target = 32768;
int64_t g_prolong_target = 65536;
// the FP64 update + residual of mixed-precision refinement: 4.59 ms with 16384
// waves, 4.48 ms with 65536 (1025 x 1025 x 513)
int64_t g_upd_target = 65536;
extern int64_t g_sweep_target_wgs;
static inline int64_t target_wgs(int64_t own) { return own > 0 ? own : g_sweep_target_wgs; }

// shortest x march of a launch over a box of `points` points: levels below
// "sweep.small_points" points are latency-bound marches whose cost grows with
// the chunk length (65 x 65 x 33 and smaller, FP64: one-pass sweep 31 us with
// chunks of 4 planes, 17 us with 1), so they use "sweep.small_min_chunk"
int64_t g_small_points = 1 << 18;
int64_t g_small_min_chunk = 1;
// the Gauss-Seidel smoother pairs its sweeps (mgp_rbgs_sweep2_*) on levels of
// at least this many points; below it two one-pass sweeps are faster (FP64:
// 257 x 257 x 129 pair 228 us, two sweeps 169 us; 513 x 513 x 257 pair
// 0.83 ms, two sweeps 0.94 ms)
int64_t g_pair_min_points = 1 << 25;
extern int64_t g_sweep_min_chunk;
static inline int64_t min_chunk_for(int64_t points) {
    return points < g_small_points ? g_small_min_chunk : g_sweep_min_chunk;
}
static inline int64_t box_points(const int64_t b[6]) { return (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]); }

namespace {

constexpr int RW = 64;        // ring-1 rectangle width (z)
constexpr int RH = 16;        // ring-1 rectangle height (y)
constexpr int TZ = RW - 2;    // interior columns per tile
constexpr int TY = RH - 2;    // interior rows per tile
constexpr int UW = RW + 2;    // old-plane rectangle (2-point ring) width
constexpr int UH = RH + 2;    // ... height
constexpr int NBORDER = 2 * UW + 2 * RH;  // outermost ring of the old-plane rectangle

template <class T>
struct MarchArgs {
    View<T> ui, uo, f;
    View<T> c;  // coarse level whose prolongation the pass adds first (two-sweep pass only)
    const T* g[6];
    T two_h;
    int64_t bx0, bx1, by0, by1, bz0, bz1;
    int64_t lin0;
    T one_omega, omega_sixth, hsq;
    int ytiles, ztiles, ntiles, ntiles_pad;
    int64_t xchunk;
};

// The shape points outside the box, copied from ui to uo by extra blocks of a
// sweep launch (the blocks after the tiles): an out-of-place pass then leaves
// a complete image of what the in-place half-sweeps leave, whichever of the
// two buffers ends up current (the boundary points' own history of updates,
// signs of zero included, stays with the solution).  The halo is not copied:
// ghost values are formed again before every use.
constexpr int COPY_PER_THREAD = 8;

template <class T>
inline int64_t complement_points(const View<T>& u, const int64_t b[6]) {
    const int64_t BX = b[1] - b[0], BY = b[3] - b[2];
    return (b[0] + u.n0 - b[1]) * u.n1 * u.n2 + BX * (b[2] + u.n1 - b[3]) * u.n2 + BX * BY * (b[4] + u.n2 - b[5]);
}
template <class T>
inline int complement_blocks(const View<T>& u, const int64_t b[6], int threads) {
    const int64_t per = (int64_t)threads * COPY_PER_THREAD;
    return (int)((complement_points(u, b) + per - 1) / per);
}

// points [blk * blockDim * COPY_PER_THREAD, ...) of the complement, in the
// order: x outside the box (whole planes), then y outside (box planes), then
// z outside (box rows); consecutive threads take consecutive z where it runs.
// PRO: ui + P(coarse) instead of ui (the pass that adds the prolongation);
// ZERO: 0 (the pass whose input is zero and not read)
template <class T, bool PRO = false, bool ZERO = false>
__device__ __forceinline__ void copy_complement(const MarchArgs<T>& A, int blk) {
    const int64_t n0 = A.ui.n0, n1 = A.ui.n1, n2 = A.ui.n2;
    const int64_t BX = A.bx1 - A.bx0, BY = A.by1 - A.by0;
    const int64_t xo = A.bx0 + n0 - A.bx1, yo = A.by0 + n1 - A.by1, zo = A.bz0 + n2 - A.bz1;
    const int64_t c1 = xo * n1 * n2, c2 = BX * yo * n2, c3 = BX * BY * zo;
    const int64_t per = (int64_t)blockDim.x * COPY_PER_THREAD;
    const int64_t end = min(c1 + c2 + c3, (int64_t)(blk + 1) * per);
    for (int64_t t = (int64_t)blk * per + threadIdx.x; t < end; t += blockDim.x) {
        int64_t x, y, z;
        if (t < c1) {
            const int64_t a = t / (n1 * n2), r = t % (n1 * n2);
            x = a < A.bx0 ? a : A.bx1 + (a - A.bx0);
            y = r / n2;
            z = r % n2;
        } else if (t < c1 + c2) {
            const int64_t u = t - c1, a = u / (yo * n2), r = u % (yo * n2), c = r / n2;
            x = A.bx0 + a;
            y = c < A.by0 ? c : A.by1 + (c - A.by0);
            z = r % n2;
        } else {
            const int64_t u = t - c1 - c2, a = u / (BY * zo), r = u % (BY * zo), c = r % zo;
            x = A.bx0 + a;
            y = A.by0 + r / zo;
            z = c < A.bz0 ? c : A.bz1 + (c - A.bz0);
        }
        if (ZERO)
            A.uo.at(x, y, z) = (T)0;
        else if (PRO)
            A.uo.at(x, y, z) = A.ui.at(x, y, z) + trilinear<T>(A.c, x, y, z);
        else
            A.uo.at(x, y, z) = A.ui.at(x, y, z);
    }
}

// element at a 32-bit byte offset from a wave-uniform base: lets the compiler
// use the scalar-base + 32-bit vector-offset addressing form
template <class T>
__device__ __forceinline__ T ld(const T* base, uint32_t byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <class T>
__device__ __forceinline__ void st(T* base, uint32_t byte_off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

template <class T, int MODE>
__device__ __forceinline__ T update(T vc, T vW, T vE, T vS, T vN, T vB, T vT, T fc, T one_omega, T omega_sixth, T hsq) {
    if (MODE == MGP_MODE_REFERENCE) {
        return one_omega * vc + omega_sixth * (vW + vE + vS + vN + vB + vT - hsq * fc);
    } else {
        const T s = ((vW - vc) + (vE - vc)) + ((vS - vc) + (vN - vc)) + ((vB - vc) + (vT - vc));
        return fma(omega_sixth, fma(-hsq, fc, s), vc);
    }
}

// In-plane Neumann ghost of a ring position: sign (+-1; 0 = not a ghost),
// mirror row/column offset in the tile, face array and its element for plane 0
// and step per plane (face layouts (nx,1,nz) and (nx,ny,1)).
template <class T>
struct Ghost {
    T sign;
    int dr, dc;  // mirror = (rr + dr, cc + dc)
    const T* g;
    int64_t gstride;
};

template <class T>
__device__ __forceinline__ Ghost<T> ghost_of(const MarchArgs<T>& A, int64_t n1, int64_t n2, int64_t y, int64_t z) {
    Ghost<T> gh;
    gh.sign = 0;
    gh.dr = gh.dc = 0;
    gh.g = nullptr;
    gh.gstride = 0;
    if (z >= 0 && z < n2 && (y == -1 || y == n1) && A.g[y < 0 ? MGP_SOUTH : MGP_NORTH]) {
        gh.sign = y < 0 ? (T)-1.0 : (T)1.0;
        gh.dr = y < 0 ? 2 : -2;
        gh.g = A.g[y < 0 ? MGP_SOUTH : MGP_NORTH] + z;
        gh.gstride = n2;
    } else if (y >= 0 && y < n1 && (z == -1 || z == n2) && A.g[z < 0 ? MGP_BOTTOM : MGP_TOP]) {
        gh.sign = z < 0 ? (T)-1.0 : (T)1.0;
        gh.dc = z < 0 ? 2 : -2;
        gh.g = A.g[z < 0 ? MGP_BOTTOM : MGP_TOP] + y;
        gh.gstride = n1;
    }
    return gh;
}

// NQ ring rows per thread; RH / NQ waves per workgroup
template <class T, int MODE, int NQ>
__global__ void __launch_bounds__(64 * (RH / NQ)) k_rbgs_march(MarchArgs<T> A) {
    constexpr int WAVES = RH / NQ;
    static_assert(64 * WAVES >= NBORDER, "too few threads for the border");
    __shared__ T sU[UH][UW];
    __shared__ T sP[2][RH][RW];

    // XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin, so
    // give each XCD a contiguous run of tiles (neighbours share halo rows in
    // that XCD's L2)
    if ((int)blockIdx.x >= A.ntiles_pad) {
        copy_complement(A, (int)blockIdx.x - A.ntiles_pad);
        return;
    }
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    const int tz = tl % A.ztiles;
    const int ty = (tl / A.ztiles) % A.ytiles;
    const int xc = tl / (A.ztiles * A.ytiles);
    const int64_t y0 = A.by0 + (int64_t)ty * TY;
    const int64_t z0 = A.bz0 + (int64_t)tz * TZ;
    const int64_t xa = A.bx0 + (int64_t)xc * A.xchunk;
    const int64_t xb = min(A.bx1, xa + A.xchunk);

    const int64_t n0 = A.ui.n0, n1 = A.ui.n1, n2 = A.ui.n2, s0 = A.ui.s0, s1 = A.ui.s1;
    const T two_h = A.two_h;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int cc = lane + 1;
    const int64_t z = z0 - 1 + lane;
    const bool z_in = z >= A.bz0 && z < A.bz1;
    const bool z_shape = z >= 0 && z < n2;
    const bool z_core = lane >= 1 && lane <= TZ;

    // per owned position: element offset within a plane (clamped into the
    // shape so every load is valid), and masks
    uint32_t off[NQ], foff[NQ];  // element offset in a plane; index in an x-face array
    bool inb[NQ], core[NQ], inshape[NQ];
    int col0[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int rr = 1 + wave * NQ + q;
        const int64_t y = y0 - 2 + rr;
        inshape[q] = z_shape && y >= 0 && y < n1;
        inb[q] = z_in && y >= A.by0 && y < A.by1;
        core[q] = inb[q] && z_core && rr >= 2 && rr <= TY + 1;
        const int64_t yc = min(max(y, (int64_t)0), n1 - 1), zc = min(max(z, (int64_t)0), n2 - 1);
        off[q] = (uint32_t)((yc * s1 + zc) * (int64_t)sizeof(T));   // bytes
        foff[q] = (uint32_t)((yc * n2 + zc) * (int64_t)sizeof(T));  // bytes
        // colour at plane 0 (reference: parity of the halo-inclusive index,
        // include/gaussseidel.h:56,81; consistent: parity of i+j+k)
        col0[q] = MODE == MGP_MODE_REFERENCE ? (int)((A.lin0 + y * s1 + z) & 1) : (int)((y + z) & 1);
    }
    // border position of this thread (outermost ring of the old-plane rectangle)
    int brr = 0, bcc = 0;
    bool has_border = false;
    {
        const int b = threadIdx.x;
        if (b < UW) brr = 0, bcc = b, has_border = true;
        else if (b < 2 * UW) brr = UH - 1, bcc = b - UW, has_border = true;
        else if (b < 2 * UW + RH) brr = 1 + (b - 2 * UW), bcc = 0, has_border = true;
        else if (b < NBORDER) brr = 1 + (b - 2 * UW - RH), bcc = UW - 1, has_border = true;
    }
    const int64_t by = y0 - 2 + brr, bz = z0 - 2 + bcc;
    const bool b_shape = has_border && by >= 0 && by < n1 && bz >= 0 && bz < n2;
    const int64_t byc = min(max(by, (int64_t)0), n1 - 1), bzc = min(max(bz, (int64_t)0), n2 - 1);
    const uint32_t boff = (uint32_t)((byc * s1 + bzc) * (int64_t)sizeof(T));
    const uint32_t bfoff = (uint32_t)((byc * n2 + bzc) * (int64_t)sizeof(T));
    // Neumann ghost positions in this tile's old-plane rectangle, rows
    // y0-2 .. y0+TY+1 and columns z0-2 .. z0+TZ+1 (uniform: boundary tiles only)
    const bool yz_ghosts = (y0 - 2 < 0 && A.g[MGP_SOUTH]) || (y0 + TY + 1 >= n1 && A.g[MGP_NORTH]) ||
                           (z0 - 2 < 0 && A.g[MGP_BOTTOM]) || (z0 + TZ + 1 >= n2 && A.g[MGP_TOP]);
    const T* gW = A.g[MGP_WEST];
    const T* gE = A.g[MGP_EAST];

    // plane j of the old values at the owned positions and at the border
    // position; the x-ghost planes -1 and n0 (Neumann west/east) are
    // sign*2h*g + the mirror plane (include/neumann.h:123-124 operand order),
    // formed at the two steps that need them
    const T* up = A.ui.p;
    auto plane_of = [&](int64_t j) {
        return up + min(max(j < 0 ? (int64_t)1 : (j >= n0 ? n0 - 2 : j), (int64_t)0), n0 - 1) * s0;
    };
    auto xghost = [&](int64_t j, uint32_t fo, T v) -> T {  // j outside [0, n0): uniform
        const T* g = j < 0 ? gW : gE;
        if (!g) return v;
        const T sg = j < 0 ? (T)-1.0 : (T)1.0;
        return sg * two_h * ld(g, fo) + v;  // face (1, ny, nz): y*nz + z
    };
    auto load_owned = [&](int64_t j, T (&dst)[NQ]) {
        const T* pl = plane_of(j);
#pragma unroll
        for (int q = 0; q < NQ; q++) dst[q] = inshape[q] ? ld(pl, off[q]) : (T)0.0;
        if (j < 0 || j >= n0) {
#pragma unroll
            for (int q = 0; q < NQ; q++) dst[q] = inshape[q] ? xghost(j, foff[q], dst[q]) : (T)0.0;
        }
    };
    auto load_border = [&](int64_t j) -> T {
        const T v = b_shape ? ld(plane_of(j), boff) : (T)0.0;
        return (j < 0 || j >= n0) && b_shape ? xghost(j, bfoff, v) : v;
    };
    const T* fpl = A.f.p;
    auto load_f = [&](int64_t j, T (&dst)[NQ]) {
        const T* pl = fpl + min(max(j, (int64_t)0), n0 - 1) * s0;
#pragma unroll
        for (int q = 0; q < NQ; q++) dst[q] = ld(pl, off[q]);
    };

    T uo_m[NQ], uo_c[NQ], uo_p[NQ], uo_n[NQ], f_m[NQ], f_c[NQ], f_n[NQ], Pm2[NQ], Pm1[NQ], Pc[NQ];
    const int64_t i0 = xa - 1;
    load_owned(i0 - 1, uo_m);
    load_owned(i0, uo_c);
    load_owned(i0 + 1, uo_p);
    load_owned(i0 + 2, uo_n);
    T bd_c = load_border(i0), bd_n = load_border(i0 + 1);
    load_f(i0, f_c);
    load_f(i0 + 1, f_n);
#pragma unroll
    for (int q = 0; q < NQ; q++) f_m[q] = Pm2[q] = Pm1[q] = Pc[q] = 0;

    for (int64_t i = i0; i <= xb; i++) {
        const bool plane_in = i >= 0 && i < n0;  // uniform
        const bool x_in = i >= A.bx0 && i < A.bx1;
        const int par = (int)(i & 1);
        // A: old plane i into LDS
        if (plane_in) {
#pragma unroll
            for (int q = 0; q < NQ; q++) sU[1 + wave * NQ + q][cc] = uo_c[q];
            if (has_border) sU[brr][bcc] = bd_c;
        }
        __syncthreads();
        if (yz_ghosts && plane_in) {
            // in-plane ghosts of the old plane (owned and border positions):
            // sign*2h*g + mirror, the values the ghost refresh before the red
            // half-sweep stores
            T gv[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int rr = 1 + wave * NQ + q;
                const Ghost<T> gh = ghost_of(A, n1, n2, y0 - 2 + rr, z);
                gv[q] = gh.sign != (T)0.0 ? gh.sign * two_h * gh.g[i * gh.gstride] + sU[rr + gh.dr][cc + gh.dc]
                                          : uo_c[q];
            }
            const Ghost<T> bg = ghost_of(A, n1, n2, by, bz);
            const bool b_ghost = has_border && bg.sign != (T)0.0;
            const T bgv = b_ghost ? bg.sign * two_h * bg.g[i * bg.gstride] + sU[brr + bg.dr][bcc + bg.dc] : (T)0.0;
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int rr = 1 + wave * NQ + q;
                sU[rr][cc] = gv[q];
                uo_c[q] = gv[q];
            }
            if (b_ghost) sU[brr][bcc] = bgv;
            __syncthreads();
        }
        // B: red points of plane i on the ring rectangle, from old values
        if (plane_in) {
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int rr = 1 + wave * NQ + q;
                const T r = update<T, MODE>(uo_c[q], uo_m[q], uo_p[q], sU[rr - 1][cc], sU[rr + 1][cc],
                                            sU[rr][cc - 1], sU[rr][cc + 1], f_c[q], A.one_omega, A.omega_sixth,
                                            A.hsq);
                const bool red = x_in && inb[q] && ((col0[q] ^ par) == 0);
                Pc[q] = red ? r : uo_c[q];
                sP[par][rr - 1][lane] = Pc[q];
            }
        }
        __syncthreads();
        if (yz_ghosts && plane_in) {
            // in-plane ghosts of the new red plane (the refresh before the
            // black half-sweep)
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int rr = 1 + wave * NQ + q;
                const Ghost<T> gh = ghost_of(A, n1, n2, y0 - 2 + rr, z);
                if (gh.sign != (T)0.0) Pc[q] = gh.sign * two_h * gh.g[i * gh.gstride] + sP[par][rr - 1 + gh.dr][lane + gh.dc];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NQ; q++) sP[par][wave * NQ + q][lane] = Pc[q];
            __syncthreads();
        }
        // C: black points of plane p = i-1 on the tile interior, from the new
        // red values; both colours of plane p written out
        const int64_t p = i - 1;
        if (p >= xa && p < xb) {
            const int pb = par ^ 1;
            T* outp = A.uo.p + p * s0;
            // x-neighbours: new red planes p-1 and p+1, or their ghosts beyond
            // the x faces (uniform branches)
            const bool w_ghost = p == 0 && gW, e_ghost = p == n0 - 1 && gE;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int r0 = wave * NQ + q;
                T vW = Pm2[q], vE = Pc[q];
                if (w_ghost) vW = (T)-1.0 * two_h * ld(gW, foff[q]) + Pc[q];
                if (e_ghost) vE = (T)1.0 * two_h * ld(gE, foff[q]) + Pm2[q];
                const int rS = max(r0 - 1, 0), rN = min(r0 + 1, RH - 1);
                const int lB = max(lane - 1, 0), lT = min(lane + 1, RW - 1);
                const T b = update<T, MODE>(uo_m[q], vW, vE, sP[pb][rS][lane], sP[pb][rN][lane], sP[pb][r0][lB],
                                            sP[pb][r0][lT], f_m[q], A.one_omega, A.omega_sixth, A.hsq);
                const bool black = (col0[q] ^ pb) == 1;
                if (core[q]) st(outp, off[q], black ? b : Pm1[q]);
            }
        }
        // rotate the x window and prefetch the plane after next
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            uo_m[q] = uo_c[q];
            uo_c[q] = uo_p[q];
            uo_p[q] = uo_n[q];
            f_m[q] = f_c[q];
            f_c[q] = f_n[q];
            Pm2[q] = Pm1[q];
            Pm1[q] = Pc[q];
        }
        bd_c = bd_n;
        load_owned(i + 3, uo_n);
        bd_n = load_border(i + 2);
        load_f(i + 2, f_n);
    }
}


// ---------------------------------------------------------------------------
// Wave-private variant: one wave owns a (y, z) tile and marches along x with
// every value in registers; no LDS and no barriers, so each wave streams at
// its own pace and keeps a whole plane of loads in flight.
//
// Lane l holds column z = z0 - 2 + l; register row t (0 .. R+1) holds
// y = y0 - 2 + t.  Red points are updated on lanes 1..62 and rows 1..R (the
// tile plus a one-point ring), black points on lanes 2..61 and rows 2..R-1
// (the tile, WZ x (R-2) points), which are the ones written.  z-neighbours come
// from the adjacent lanes through DPP wave shifts, y-neighbours from the
// adjacent register rows, x-neighbours from the previous and next planes'
// registers.  A plane's registers hold the old values until its red points
// are updated in place, then the new red values (black points keep the old
// value until their own update one step later), which is exactly what each
// half-sweep's stencil reads.
constexpr int WZ = 60;  // interior columns per wave tile

// DPP lane move of a 32- or 64-bit value with bound_ctrl: a lane whose source
// lane lies outside the wave reads 0.  No old value is involved, so the
// compiler emits no initialising move before each DPP move.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    typedef int i32x2 __attribute__((ext_vector_type(2)));
    i32x2 w = __builtin_bit_cast(i32x2, v);
    w.x = __builtin_amdgcn_mov_dpp(w.x, CTRL, 0xf, 0xf, true);
    w.y = __builtin_amdgcn_mov_dpp(w.y, CTRL, 0xf, 0xf, true);
    return __builtin_bit_cast(double, w);
}
template <class T>
__device__ __forceinline__ T lane_prev(T v) {  // value of lane l-1 (z - 1); 0 on lane 0
    return dpp_move<0x138>(v);                 // wave_shr:1
}
template <class T>
__device__ __forceinline__ T lane_next(T v) {  // value of lane l+1 (z + 1); 0 on lane 63
    return dpp_move<0x130>(v);                 // wave_shl:1
}

// Buffer-resource access (SGPR descriptor + 32-bit lane offset + SGPR row
// offset): no per-lane address arithmetic.  A lane offset at or beyond the
// record count drops the store in hardware, which masks stores without
// branches; only the lane (VGPR) offset is used for that, since the range
// check does not cover the SGPR offset.  Loads always use in-range offsets.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t BUF_OOB = 0x40000000u;  // offset past any plane (planes < 1 GiB)
// cache-policy bits of the buffer loads and stores (gfx950: bit 0 sc0, bit 1
// nt, bit 4 sc1); a build-time choice, measured per kernel
#ifndef MGP_LOAD_POLICY
#define MGP_LOAD_POLICY 0
#endif
#ifndef MGP_STORE_POLICY
#define MGP_STORE_POLICY 0
#endif

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
template <class T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, MGP_LOAD_POLICY));
}
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, MGP_LOAD_POLICY));
}
__device__ __forceinline__ void bstore(double v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)voff, (int)soff, MGP_STORE_POLICY);
}
__device__ __forceinline__ void bstore(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)voff, (int)soff,
                                          MGP_STORE_POLICY);
}

// EDGE: the tile has rows outside the box or in-plane Neumann ghost positions
// (boundary tiles only); the other tiles run a copy of the loop without the
// row masks and the ghost code
// D: planes of loads in flight ahead of the red update (1 or 2)
template <class T, int MODE, int R, int D, bool EDGE>
__device__ __forceinline__ void wave_march(const MarchArgs<T>& A, int tl) {
    constexpr int NR = R + 2;
    const int tz = tl % A.ztiles;
    const int ty = (tl / A.ztiles) % A.ytiles;
    const int xc = tl / (A.ztiles * A.ytiles);
    const int64_t y0 = A.by0 + (int64_t)ty * (R - 2);
    const int64_t z0 = A.bz0 + (int64_t)tz * WZ;
    const int64_t xa = A.bx0 + (int64_t)xc * A.xchunk;
    const int64_t xb = min(A.bx1, xa + A.xchunk);
    const int64_t n0 = A.ui.n0, n1 = A.ui.n1, n2 = A.ui.n2, s0 = A.ui.s0, s1 = A.ui.s1;
    const T two_h = A.two_h;
    // bytes from a plane's (y, z) = (0, 0) to the end of its last row
    const uint32_t plane_bytes = (uint32_t)(((n1 - 1) * s1 + n2) * (int64_t)sizeof(T));

    const int lane = threadIdx.x;
    const int64_t z = z0 - 2 + lane;
    const bool z_in = z >= A.bz0 && z < A.bz1;
    const bool z_shape = z >= 0 && z < n2;
    const bool lane_red = z_in && lane >= 1 && lane <= 62;
    const bool lane_core = z_in && lane >= 2 && lane <= 61;
    const int64_t zc = min(max(z, (int64_t)0), n2 - 1);
    const uint32_t zoff = (uint32_t)(zc * (int64_t)sizeof(T));
    const uint32_t zoff_st = lane_core ? zoff : BUF_OOB;
    // colour: red where (i + y + z) is even (consistent), or where the
    // halo-inclusive linear index is even (reference, include/gaussseidel.h:56,81;
    // odd strides make that the parity of lin0 + i + y + z)
    const int cbase = (int)(((MODE == MGP_MODE_REFERENCE ? A.lin0 : 0) + z + y0) & 1);

    // row byte offsets, clamped into the shape, and the rows inside the box
    uint32_t roff[NR];
    bool ybox[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) {
        const int64_t y = y0 - 2 + t;
        roff[t] = (uint32_t)(min(max(y, (int64_t)0), n1 - 1) * s1 * (int64_t)sizeof(T));
        ybox[t] = y >= A.by0 && y < A.by1;
    }

    const T* gW = A.g[MGP_WEST];
    const T* gE = A.g[MGP_EAST];
    const T* gS = A.g[MGP_SOUTH];
    const T* gN = A.g[MGP_NORTH];
    const T* gB = A.g[MGP_BOTTOM];
    const T* gT = A.g[MGP_TOP];
    const bool y_ghosts = (y0 - 2 < 0 && gS) || (y0 + R - 1 >= n1 && gN);
    const bool z_ghosts = (z0 - 2 < 0 && gB) || (z0 + WZ + 1 >= n2 && gT);

    // old values of plane j (rows 0..R+1); planes -1 and n0 are the Neumann
    // x-ghosts sign*2h*g + mirror plane (include/neumann.h:123-124), formed
    // from the old mirror plane
    auto load_plane = [&](int64_t j, T (&P)[NR]) {
        const int64_t jm = min(max(j < 0 ? (int64_t)1 : (j >= n0 ? n0 - 2 : j), (int64_t)0), n0 - 1);
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(A.ui.p + jm * s0, plane_bytes);
#pragma unroll
        for (int t = 0; t < NR; t++) P[t] = bload<T>(r, zoff, roff[t]);
        if ((j < 0 && gW) || (j >= n0 && gE)) {
            const T* g = j < 0 ? gW : gE;
            const T sg = j < 0 ? (T)-1.0 : (T)1.0;
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t y = min(max(y0 - 2 + t, (int64_t)0), n1 - 1);
                P[t] = sg * two_h * g[y * n2 + zc] + P[t];  // face (1, ny, nz)
            }
        }
    };
    auto load_f = [&](int64_t j, T (&F)[NR]) {
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(A.f.p + min(max(j, (int64_t)0), n0 - 1) * s0, plane_bytes);
#pragma unroll
        for (int t = 1; t <= R; t++) F[t] = bload<T>(r, zoff, roff[t]);
    };
    // in-plane Neumann ghosts of plane i from the values in P (the ghost
    // refresh before a half-sweep): y ghosts from the rows two apart, z ghosts
    // from the lanes two apart
    auto fix_ghosts = [&](int64_t i, T (&P)[NR]) {
        if (!EDGE) return;
        if (y_ghosts) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t y = y0 - 2 + t;
                if (t + 2 < NR && y == -1 && gS) {
                    const T v = (T)-1.0 * two_h * gS[i * n2 + zc] + P[t + 2];  // face (nx, 1, nz)
                    P[t] = z_shape ? v : P[t];
                } else if (t >= 2 && y == n1 && gN) {
                    const T v = (T)1.0 * two_h * gN[i * n2 + zc] + P[t - 2];
                    P[t] = z_shape ? v : P[t];
                }
            }
        }
        if (z_ghosts) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t y = y0 - 2 + t;
                if (y < 0 || y >= n1) continue;
                if (gB) {
                    const T m = lane_next(lane_next(P[t]));
                    const T v = (T)-1.0 * two_h * gB[i * n1 + y] + m;  // face (nx, ny, 1)
                    P[t] = z == -1 ? v : P[t];
                }
                if (gT) {
                    const T m = lane_prev(lane_prev(P[t]));
                    const T v = (T)1.0 * two_h * gT[i * n1 + y] + m;
                    P[t] = z == n2 ? v : P[t];
                }
            }
        }
    };

    // x window: planes i-2 .. i+1 (+ i+2 when D == 2) and the plane being
    // loaded; f of planes i-1, i (+ i+1) and the f plane being loaded
    T Pm2[NR], Pm1[NR], P0[NR], Pp1[NR], Pp2[NR], PL[NR], Fm1[NR], F0[NR], Fp1[NR], FL[NR];
    const int64_t i0 = xa - 1;
    load_plane(i0 - 1, Pm1);
    load_plane(i0, P0);
    load_plane(i0 + 1, Pp1);
    load_f(i0, F0);
    if (D == 2) {
        load_plane(i0 + 2, Pp2);
        load_f(i0 + 1, Fp1);
    }
#pragma unroll
    for (int t = 0; t < NR; t++) Pm2[t] = Fm1[t] = (T)0;

    for (int64_t i = i0; i <= xb; i++) {
        // next planes in flight during this step (and the next, when D == 2)
        load_plane(i + 1 + D, PL);
        load_f(i + D, FL);
        const bool plane_in = i >= 0 && i < n0;  // uniform
        const bool x_in = i >= A.bx0 && i < A.bx1;
        const int par = (int)(i & 1);
        if (plane_in) fix_ghosts(i, P0);
        // red points of plane i
        if (x_in) {
#pragma unroll
            for (int t = 1; t <= R; t++) {
                const bool red = lane_red && (!EDGE || ybox[t]) && ((cbase ^ par ^ (t & 1)) == 0);
                const T r = update<T, MODE>(P0[t], Pm1[t], Pp1[t], P0[t - 1], P0[t + 1], lane_prev(P0[t]),
                                            lane_next(P0[t]), F0[t], A.one_omega, A.omega_sixth, A.hsq);
                P0[t] = red ? r : P0[t];
            }
        }
        if (plane_in) fix_ghosts(i, P0);
        // black points of plane p = i-1, then both colours of plane p out
        const int64_t p = i - 1;
        if (p >= xa && p < xb) {
            const int pb = par ^ 1;
            const __amdgpu_buffer_rsrc_t ro = plane_rsrc(A.uo.p + p * s0, plane_bytes);
            // x-ghost planes beyond Neumann x faces, from the new red values of
            // the mirror plane (the ghost refresh before the black half-sweep);
            // planes -1 and n0 are not needed after this step (n0 >= 3)
            if (p == 0 && gW) {
#pragma unroll
                for (int t = 2; t < R; t++) {
                    const int64_t yc = min(max(y0 - 2 + t, (int64_t)0), n1 - 1);
                    Pm2[t] = (T)-1.0 * two_h * gW[yc * n2 + zc] + P0[t];
                }
            }
            if (p == n0 - 1 && gE) {
#pragma unroll
                for (int t = 2; t < R; t++) {
                    const int64_t yc = min(max(y0 - 2 + t, (int64_t)0), n1 - 1);
                    P0[t] = (T)1.0 * two_h * gE[yc * n2 + zc] + Pm2[t];
                }
            }
#pragma unroll
            for (int t = 2; t < R; t++) {
                const bool black = lane_core && ((cbase ^ pb ^ (t & 1)) == 1);
                const T b = update<T, MODE>(Pm1[t], Pm2[t], P0[t], Pm1[t - 1], Pm1[t + 1], lane_prev(Pm1[t]),
                                            lane_next(Pm1[t]), Fm1[t], A.one_omega, A.omega_sixth, A.hsq);
                Pm1[t] = black ? b : Pm1[t];
                bstore(Pm1[t], ro, !EDGE || ybox[t] ? zoff_st : BUF_OOB, roff[t]);
            }
        }
        // rotate the x window
#pragma unroll
        for (int t = 0; t < NR; t++) {
            Pm2[t] = Pm1[t];
            Pm1[t] = P0[t];
            P0[t] = Pp1[t];
            Fm1[t] = F0[t];
            if (D == 2) {
                Pp1[t] = Pp2[t];
                Pp2[t] = PL[t];
                F0[t] = Fp1[t];
                Fp1[t] = FL[t];
            } else {
                Pp1[t] = PL[t];
                F0[t] = FL[t];
            }
        }
    }
}

template <class T, int MODE, int R, int D>
__global__ void __launch_bounds__(64) k_rbgs_wave(MarchArgs<T> A) {
    if ((int)blockIdx.x >= A.ntiles_pad) {
        copy_complement(A, (int)blockIdx.x - A.ntiles_pad);
        return;
    }
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    const int ty = (tl / A.ztiles) % A.ytiles, tz = tl % A.ztiles;
    const int64_t y0 = A.by0 + (int64_t)ty * (R - 2), z0 = A.bz0 + (int64_t)tz * WZ;
    // red rows y0-1 .. y0+R-2 all inside the box, and no ghost positions
    const bool edge = y0 - 1 < A.by0 || y0 + R - 2 >= A.by1 || (y0 - 2 < 0 && A.g[MGP_SOUTH]) ||
                      (y0 + R - 1 >= A.ui.n1 && A.g[MGP_NORTH]) || (z0 - 2 < 0 && A.g[MGP_BOTTOM]) ||
                      (z0 + WZ + 1 >= A.ui.n2 && A.g[MGP_TOP]);
    if (edge) wave_march<T, MODE, R, D, true>(A, tl);
    else wave_march<T, MODE, R, D, false>(A, tl);
}


// ---------------------------------------------------------------------------
// Two sweeps in one pass (mgp_rbgs_sweep2_*): the wave march above with four
// half-sweep stages in flight instead of two.  At plane step i the wave updates
//   stage 0: sweep 1 red   of plane i,
//   stage 1: sweep 1 black of plane i-1,
//   stage 2: sweep 2 red   of plane i-2,
//   stage 3: sweep 2 black of plane i-3, then writes plane i-3,
// in that order, so each stage's x-neighbours hold exactly the values the
// previous half-sweep left there: the planes on both sides of a stage's plane
// have been through the previous stage and not yet through this one.  In y and
// z every stage is valid one point further inside the tile than the stage
// before it: register rows t = 0 .. NR-1 (y = y0 - 4 + t) and lanes 0 .. 63
// (z = z0 - 4 + lane) shrink to rows 4 .. NR-5 and lanes 4 .. 59 after stage
// 3, which is what the tile writes (NR - 8 rows x 56 columns; the rings are
// recomputed by the neighbouring tiles).  Neumann ghosts are formed before
// every stage from the values the stage reads (the ghost refresh before each
// half-sweep, include/gaussseidel.h:28-94 with include/neumann.h:108-124):
// in-plane ghosts from the rows / lanes two apart, the x-ghost planes -1 and
// n0 from the planes two apart.  Per point the arithmetic is the single
// sweep's, so the result equals two sweeps bit for bit.
//
// PRO (mgp_rbgs_sweep2_prolong_*): the pass first adds the trilinear
// prolongation of a coarse level to every plane as it arrives, so the cycle's
// correction u += P(u_coarse) (include/vcycle.h:81-94) and the two sweeps after
// it take one pass over u.  The correction is formed with the operations of
// mgp_prolong_add_* (x-, then y-, then z-midpoints, each 0.5 * (left + right),
// include/trilinearinterpolation.h:50-133), so the values are the same.  Each
// lane keeps one coarse column q at rows cy0 .. cy0 + NR/2 of the two coarse
// planes under the fine plane j (j >> 1 and (j >> 1) + 1, in LDS); an odd z
// takes its second column from the neighbouring lane.  PY is the parity of y at
// register row 0 and PZ that of z at lane 0 (the launch fixes both: tiles step
// by even counts): PZ = 0 has each lane read column (z + 1) >> 1, so an odd z
// finds column z >> 1 on the lane before; PZ = 1 has it read z >> 1 and find
// (z >> 1) + 1 on the lane after.  Rows, lanes and planes outside the shape
// receive values that no update reads (they are masked, or ghosts formed
// afresh before every stage).
//
// ZIN (mgp_rbgs_sweep2_from_zero_*): the input planes are zeros and are not
// read; the passes' arithmetic is unchanged, so the result is that of a fill of
// 0 followed by the pass.  PACK: the packed stages (stage lambda below).
constexpr int WZ2 = 56;  // interior columns per two-sweep wave tile

template <class T, int MODE, int NR, bool EDGE, bool PRO = false, int PY = 0, int PZ = 0, bool PACK = false,
          bool ZIN = false, bool FLDS = false>
__device__ __forceinline__ void wave_march2(const MarchArgs<T>& A, int tl, T* cl = nullptr, T* fr = nullptr) {
    // FLDS (with PACK): the f planes in the wave's LDS (fr: a ring of 4 planes x
    // NR-2 rows x 64 lanes, plane q in slot q & 3) instead of registers, each
    // lane reading the row its packed pair takes.  1025 x 1025 x 513 with the
    // correction inside: FP32 2.43 ms against 3.40 ms; FP64 5.18 ms either way;
    // the plain passes are 1 % faster with f in registers
    constexpr int RO = NR - 8;  // rows written per tile
    const int tz = tl % A.ztiles;
    const int ty = (tl / A.ztiles) % A.ytiles;
    const int xc = tl / (A.ztiles * A.ytiles);
    const int64_t y0 = A.by0 + (int64_t)ty * RO;
    const int64_t z0 = A.bz0 + (int64_t)tz * WZ2;
    const int64_t xa = A.bx0 + (int64_t)xc * A.xchunk;
    const int64_t xb = min(A.bx1, xa + A.xchunk);
    const int64_t n0 = A.ui.n0, n1 = A.ui.n1, n2 = A.ui.n2, s0 = A.ui.s0, s1 = A.ui.s1;
    const T two_h = A.two_h;
    const uint32_t plane_bytes = (uint32_t)(((n1 - 1) * s1 + n2) * (int64_t)sizeof(T));

    const int lane = threadIdx.x;
    const int64_t z = z0 - 4 + lane;
    const bool z_in = z >= A.bz0 && z < A.bz1;
    const bool z_shape = z >= 0 && z < n2;
    const int64_t zc = min(max(z, (int64_t)0), n2 - 1);
    const uint32_t zoff = (uint32_t)(zc * (int64_t)sizeof(T));
    const bool lane_core = z_in && lane >= 4 && lane <= 59;
    const uint32_t zoff_st = lane_core ? zoff : BUF_OOB;
    // colour as in the single sweep (rows are y0 - 4 + t: same parity as y0 + t)
    const int cbase = (int)(((MODE == MGP_MODE_REFERENCE ? A.lin0 : 0) + z + y0) & 1);

    uint32_t roff[NR];
    bool ybox[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) {
        const int64_t y = y0 - 4 + t;
        roff[t] = (uint32_t)(min(max(y, (int64_t)0), n1 - 1) * s1 * (int64_t)sizeof(T));
        ybox[t] = y >= A.by0 && y < A.by1;
    }

    const T* gW = A.g[MGP_WEST];
    const T* gE = A.g[MGP_EAST];
    const T* gS = A.g[MGP_SOUTH];
    const T* gN = A.g[MGP_NORTH];
    const T* gB = A.g[MGP_BOTTOM];
    const T* gT = A.g[MGP_TOP];
    const bool y_ghosts = (y0 - 4 < 0 && gS) || (y0 + NR - 5 >= n1 && gN);
    const bool z_ghosts = (z0 - 4 < 0 && gB) || (z0 + 59 >= n2 && gT);

    auto load_plane = [&](int64_t j, T (&P)[NR]) {
        if constexpr (ZIN) {  // an input of zeros, not read
#pragma unroll
            for (int t = 0; t < NR; t++) P[t] = (T)0;
            return;
        }
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(A.ui.p + min(max(j, (int64_t)0), n0 - 1) * s0, plane_bytes);
#pragma unroll
        for (int t = 0; t < NR; t++) P[t] = bload<T>(r, zoff, roff[t]);
    };
    auto load_f = [&](int64_t j, T (&F)[NR]) {
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(A.f.p + min(max(j, (int64_t)0), n0 - 1) * s0, plane_bytes);
#pragma unroll
        for (int t = 1; t <= NR - 2; t++) F[t] = bload<T>(r, zoff, roff[t]);
    };

    // PRO: coarse column q (clamped into the coarse shape) of rows cy0 + b;
    // the row goes into the lane offset, so rows outside the coarse shape
    // (edge tiles only; no update reads what they feed) fall outside the
    // plane's record range and read 0
    constexpr int CR = NR / 2 + 1;
    // The two coarse planes live in the wave's LDS (cl, slots of CR x 64
    // values: plane j >> 1 in slot cs, (j >> 1) + 1 in slot cs ^ 1); a new
    // plane is loaded into registers (Ct) at the start of an even fine plane's
    // step and stored to its slot at the end of the step, when it has arrived
    T Ct[CR];
    int cs = 0;
    bool ct_pending = false;
    int32_t cvoff = 0, crow = 0;
    uint32_t cplane = 0;
    if constexpr (PRO) {
        const int64_t c1 = A.c.n1, c2 = A.c.n2;
        const int64_t cy0 = (y0 - 4 - PY) >> 1;
        const int64_t q = min(max((z + 1 - PZ) >> 1, (int64_t)0), c2 - 1);
        crow = (int32_t)(A.c.s1 * (int64_t)sizeof(T));
        cvoff = (int32_t)(cy0 * crow + q * (int64_t)sizeof(T));
        cplane = (uint32_t)(((c1 - 1) * A.c.s1 + c2) * (int64_t)sizeof(T));
    }
    auto cl_at = [&](int slot, int b) -> T& { return cl[(slot * CR + b) * 64 + lane]; };
    auto load_coarse = [&](int64_t a, T (&V)[CR]) {
        const __amdgpu_buffer_rsrc_t r =
            plane_rsrc(A.c.p + min(max(a, (int64_t)0), A.c.n0 - 1) * A.c.s0, cplane);
#pragma unroll
        for (int b = 0; b < CR; b++) V[b] = bload<T>(r, (uint32_t)(cvoff + b * crow), 0);
    };
    auto store_coarse = [&](int slot, const T (&V)[CR]) {
#pragma unroll
        for (int b = 0; b < CR; b++) cl_at(slot, b) = V[b];
    };
    // the slots for fine plane j, given them for plane j - 1
    auto next_coarse = [&](int64_t j) {
        if (j & 1) return;
        cs ^= 1;  // plane (j >> 1) was the second plane of j - 1
        load_coarse((j >> 1) + 1, Ct);
        ct_pending = true;
    };
    // P[t] += P(coarse) from the x-stage values X (coarse rows cy0 + b): the
    // y-midpoint on odd rows, then the z-midpoint 0.5 * (Y + o) with o the
    // other column's Y on an odd z and Y itself on an even z (0.5 * (Y + Y) is
    // Y; a quad permutation fetches both cases in one move: PZ = 1 pairs lanes
    // 2k <- 2k+1, PZ = 0 lanes 2k+1 <- 2k)
    auto add_rows = [&](const T (&X)[CR], T (&P)[NR]) {
#pragma unroll
        for (int t = 0; t < NR; t++) {
            const int b = (t + PY) >> 1;
            const T Y = ((t + PY) & 1) ? half_sum<T>(X[b], X[b + 1 < CR ? b + 1 : b]) : X[b];
            const T o = dpp_move<PZ ? 0xF5 : 0xA0>(Y);  // quad_perm [1,1,3,3] / [0,0,2,2]
            P[t] = P[t] + half_sum<T>(Y, o);
        }
    };
    // P[t] += P(coarse) at (j, y0 - 4 + t, z)
    auto add_prolong = [&](int64_t j, T (&P)[NR]) {
        if (ct_pending) {
            store_coarse(cs ^ 1, Ct);
            ct_pending = false;
        }
        T X[CR];
        if (j & 1) {
#pragma unroll
            for (int b = 0; b < CR; b++) X[b] = half_sum<T>(cl_at(cs, b), cl_at(cs ^ 1, b));
        } else {
#pragma unroll
            for (int b = 0; b < CR; b++) X[b] = cl_at(cs, b);
        }
        add_rows(X, P);
    };

    // stage S at plane p (S even: red, odd: black) with west / centre / east
    // planes W, C, E and f of plane p; rows S+1 .. NR-2-S are updated, rows
    // S .. NR-1-S read
    auto fr_at = [&](int64_t q, int row) -> T& { return fr[(((int)q & 3) * (NR - 2) + row - 1) * 64 + lane]; };
    auto stage = [&](auto Sc, int64_t p, T (&W)[NR], T (&C)[NR], T (&E)[NR], const T (&F)[NR]) {
        constexpr int S = decltype(Sc)::value;
        // x-ghost planes from the plane two apart (face layout (1, ny, nz))
        if (p == 0 && gW) {
#pragma unroll
            for (int t = S + 1; t <= NR - 2 - S; t++) {
                const int64_t yc = min(max(y0 - 4 + t, (int64_t)0), n1 - 1);
                W[t] = (T)-1.0 * two_h * gW[yc * n2 + zc] + E[t];
            }
        }
        if (p == n0 - 1 && gE) {
#pragma unroll
            for (int t = S + 1; t <= NR - 2 - S; t++) {
                const int64_t yc = min(max(y0 - 4 + t, (int64_t)0), n1 - 1);
                E[t] = (T)1.0 * two_h * gE[yc * n2 + zc] + W[t];
            }
        }
        // in-plane ghosts of plane p among the rows the stage reads
        if (EDGE && y_ghosts) {
#pragma unroll
            for (int t = S; t <= NR - 1 - S; t++) {
                const int64_t y = y0 - 4 + t;
                if (t + 2 <= NR - 1 - S && y == -1 && gS) {
                    const T v = (T)-1.0 * two_h * gS[p * n2 + zc] + C[t + 2];  // face (nx, 1, nz)
                    C[t] = z_shape ? v : C[t];
                } else if (t - 2 >= S && y == n1 && gN) {
                    const T v = (T)1.0 * two_h * gN[p * n2 + zc] + C[t - 2];
                    C[t] = z_shape ? v : C[t];
                }
            }
        }
        if (EDGE && z_ghosts) {
#pragma unroll
            for (int t = S; t <= NR - 1 - S; t++) {
                const int64_t y = y0 - 4 + t;
                if (y < 0 || y >= n1) continue;
                if (gB) {
                    const T m = lane_next(lane_next(C[t]));
                    const T v = (T)-1.0 * two_h * gB[p * n1 + y] + m;  // face (nx, ny, 1)
                    C[t] = z == -1 ? v : C[t];
                }
                if (gT) {
                    const T m = lane_prev(lane_prev(C[t]));
                    const T v = (T)1.0 * two_h * gT[p * n1 + y] + m;
                    C[t] = z == n2 ? v : C[t];
                }
            }
        }
        const bool lane_ok = z_in && lane >= S + 1 && lane <= 62 - S;
        const int want = S & 1;  // 0: red, 1: black
        const int par = (int)(p & 1);
        if constexpr (PACK) {
            // rows in pairs (t, t+1): the lanes of the stage's colour in row t
            // (m) take row t, the others row t+1, so one stencil evaluation
            // serves both rows instead of two evaluations that each keep half
            // their lanes; the z-neighbours of either come from the
            // complementary selection x shifted by one lane (adjacent lanes
            // always differ in colour).  Same operands in the same order per
            // point, so the same bits; each pair costs selects (32-bit moves)
            // where it saves a whole FP64 stencil
            const bool m = (cbase ^ par ^ ((S + 1) & 1)) == want;
#pragma unroll
            for (int t = S + 1; t <= NR - 2 - S; t += 2) {
                const T c = m ? C[t] : C[t + 1];
                const T x = m ? C[t + 1] : C[t];
                const T v = update<T, MODE>(c, m ? W[t] : W[t + 1], m ? E[t] : E[t + 1], m ? C[t - 1] : C[t],
                                            m ? C[t + 1] : C[t + 2], lane_prev(x), lane_next(x),
                                            FLDS ? fr_at(p, m ? t : t + 1) : (m ? F[t] : F[t + 1]), A.one_omega,
                                            A.omega_sixth, A.hsq);
                const bool ok0 = lane_ok && (!EDGE || ybox[t]), ok1 = lane_ok && (!EDGE || ybox[t + 1]);
                C[t] = m && ok0 ? v : C[t];
                C[t + 1] = !m && ok1 ? v : C[t + 1];
            }
        } else {
#pragma unroll
            for (int t = S + 1; t <= NR - 2 - S; t++) {
                const bool upd = lane_ok && (!EDGE || ybox[t]) && ((cbase ^ par ^ (t & 1)) == want);
                const T v = update<T, MODE>(C[t], W[t], E[t], C[t - 1], C[t + 1], lane_prev(C[t]), lane_next(C[t]),
                                            F[t], A.one_omega, A.omega_sixth, A.hsq);
                C[t] = upd ? v : C[t];
            }
        }
    };

    // x window: planes i-4 .. i+1 and the plane being loaded (i+2); f of
    // planes i-3 .. i and the f plane being loaded (i+1)
    T Pm4[NR], Pm3[NR], Pm2[NR], Pm1[NR], P0[NR], Pp1[NR], PL[NR];
    T Fm3[NR], Fm2[NR], Fm1[NR], F0[NR], FL[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) Pm4[t] = Pm3[t] = Pm2[t] = Fm3[t] = Fm2[t] = Fm1[t] = (T)0;
    const int64_t i0 = xa - 3;
    load_plane(i0 - 1, Pm1);
    load_plane(i0, P0);
    load_plane(i0 + 1, Pp1);
    load_f(i0, F0);
    if constexpr (FLDS) {
#pragma unroll
        for (int t = 1; t <= NR - 2; t++) fr_at(i0, t) = F0[t];
    }
    if constexpr (PRO) {
        const int64_t j0 = i0 - 1;
        load_coarse(j0 >> 1, Ct);
        store_coarse(0, Ct);
        load_coarse((j0 >> 1) + 1, Ct);
        store_coarse(1, Ct);
        add_prolong(j0, Pm1);
        next_coarse(j0 + 1);
        add_prolong(j0 + 1, P0);
        next_coarse(j0 + 2);
        add_prolong(j0 + 2, Pp1);
    }
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    using S2 = std::integral_constant<int, 2>;
    using S3 = std::integral_constant<int, 3>;
    for (int64_t i = i0; i <= xb + 2; i++) {
        load_plane(i + 2, PL);
        load_f(i + 1, FL);
        if constexpr (PRO) next_coarse(i + 2);
        // each stage from the first plane the chunk's output depends on
        if (i >= A.bx0 && i < A.bx1) stage(S0{}, i, Pm1, P0, Pp1, F0);
        if (i - 1 >= max(A.bx0, xa - 2) && i - 1 < A.bx1) stage(S1{}, i - 1, Pm2, Pm1, P0, Fm1);
        if (i - 2 >= max(A.bx0, xa - 1) && i - 2 < A.bx1) stage(S2{}, i - 2, Pm3, Pm2, Pm1, Fm2);
        const int64_t p = i - 3;
        if (p >= xa && p < xb) {
            stage(S3{}, p, Pm4, Pm3, Pm2, Fm3);
            const __amdgpu_buffer_rsrc_t ro = plane_rsrc(A.uo.p + p * s0, plane_bytes);
#pragma unroll
            for (int t = 4; t <= NR - 5; t++) bstore(Pm3[t], ro, !EDGE || ybox[t] ? zoff_st : BUF_OOB, roff[t]);
        }
        if constexpr (PRO) add_prolong(i + 2, PL);
        if constexpr (FLDS) {  // f of plane i+1 into the slot plane i-3 used
#pragma unroll
            for (int t = 1; t <= NR - 2; t++) fr_at(i + 1, t) = FL[t];
        }
#pragma unroll
        for (int t = 0; t < NR; t++) {
            Pm4[t] = Pm3[t];
            Pm3[t] = Pm2[t];
            Pm2[t] = Pm1[t];
            Pm1[t] = P0[t];
            P0[t] = Pp1[t];
            Pp1[t] = PL[t];
            Fm3[t] = Fm2[t];
            Fm2[t] = Fm1[t];
            Fm1[t] = F0[t];
            F0[t] = FL[t];
        }
    }
}

// rows y0-3 .. y0+NR-6 (the ones stage 0 updates) inside the box, and no
// ghost positions among the tile's rows and lanes: false for interior tiles
template <class T, int NR>
__device__ __forceinline__ bool sweep2_edge_tile(const MarchArgs<T>& A, int tl) {
    const int ty = (tl / A.ztiles) % A.ytiles, tz = tl % A.ztiles;
    const int64_t y0 = A.by0 + (int64_t)ty * (NR - 8), z0 = A.bz0 + (int64_t)tz * WZ2;
    return y0 - 3 < A.by0 || y0 + NR - 6 >= A.by1 || (y0 - 4 < 0 && A.g[MGP_SOUTH]) ||
           (y0 + NR - 5 >= A.ui.n1 && A.g[MGP_NORTH]) || (z0 - 4 < 0 && A.g[MGP_BOTTOM]) ||
           (z0 + 59 >= A.ui.n2 && A.g[MGP_TOP]);
}

template <class T, int MODE, int NR, bool PACK = false, bool FLDS = false>
__global__ void __launch_bounds__(64) k_rbgs2_wave(MarchArgs<T> A) {
    if ((int)blockIdx.x >= A.ntiles_pad) {
        copy_complement(A, (int)blockIdx.x - A.ntiles_pad);
        return;
    }
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    __shared__ T fr[FLDS ? 4 * (NR - 2) * 64 : 1];
    if (sweep2_edge_tile<T, NR>(A, tl))
        wave_march2<T, MODE, NR, true, false, 0, 0, PACK, false, FLDS>(A, tl, nullptr, fr);
    else
        wave_march2<T, MODE, NR, false, false, 0, 0, PACK, false, FLDS>(A, tl, nullptr, fr);
}

// packed stages, f in LDS, held to two waves per SIMD (consistent mode; the
// row counts other than 16)
template <class T, int NR>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) k_rbgs2f_wave(MarchArgs<T> A) {
    if ((int)blockIdx.x >= A.ntiles_pad) {
        copy_complement(A, (int)blockIdx.x - A.ntiles_pad);
        return;
    }
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    __shared__ T fr[4 * (NR - 2) * 64];
    if (sweep2_edge_tile<T, NR>(A, tl))
        wave_march2<T, MGP_MODE_CONSISTENT, NR, true, false, 0, 0, true, false, true>(A, tl, nullptr, fr);
    else
        wave_march2<T, MGP_MODE_CONSISTENT, NR, false, false, 0, 0, true, false, true>(A, tl, nullptr, fr);
}

// the two sweeps of u = 0 (consistent mode): u_in is not read, the shape
// points outside the box are written 0
template <class T, int NR, bool PACK>
__global__ void __launch_bounds__(64) k_rbgs2z_wave(MarchArgs<T> A) {
    if ((int)blockIdx.x >= A.ntiles_pad) {
        copy_complement<T, false, true>(A, (int)blockIdx.x - A.ntiles_pad);
        return;
    }
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    if (sweep2_edge_tile<T, NR>(A, tl))
        wave_march2<T, MGP_MODE_CONSISTENT, NR, true, false, 0, 0, PACK, true>(A, tl);
    else
        wave_march2<T, MGP_MODE_CONSISTENT, NR, false, false, 0, 0, PACK, true>(A, tl);
}

// the two sweeps after u_in + P(coarse) (consistent mode); PY, PZ: parities of
// by0 and bz0
template <class T, int NR, int PY, int PZ, bool PACK = false>
__global__ void __launch_bounds__(64) k_rbgs2p_wave(MarchArgs<T> A) {
    if ((int)blockIdx.x >= A.ntiles_pad) {
        copy_complement<T, true>(A, (int)blockIdx.x - A.ntiles_pad);
        return;
    }
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    __shared__ T cl[2 * (NR / 2 + 1) * 64];          // the wave's two coarse planes
    __shared__ T fr[PACK ? 4 * (NR - 2) * 64 : 1];  // its f planes (packed stages)
    if (sweep2_edge_tile<T, NR>(A, tl))
        wave_march2<T, MGP_MODE_CONSISTENT, NR, true, true, PY, PZ, PACK, false, PACK>(A, tl, cl, fr);
    else
        wave_march2<T, MGP_MODE_CONSISTENT, NR, false, true, PY, PZ, PACK, false, PACK>(A, tl, cl, fr);
}


// ---------------------------------------------------------------------------
// Residual r = f - L u on a box (include/residual.h:58-79), one wave per
// (y, z) tile marching along x like the sweep: register rows for y, DPP lane
// shifts for z, the previous and next planes' registers for x.  The ghosts
// are read from the halo, which the caller has refreshed
// (include/residual.h:26-34).  With `norm`, max |r| over the box is folded into
// *norm (atomic max on the bits, exact); the caller adds the shape points
// outside the box.
template <class T, class O = T>
struct ResidArgs {
    View<T> u, f, r;
    View<O> out;  // where r goes (r itself, or a narrower copy)
    int64_t bx0, bx1, by0, by1, bz0, bz1;
    int64_t ylo, yhi, zlo, zhi, xlo, xhi;  // allocated (halo-inclusive) index ranges
    T hsq, inv_h2;
    unsigned long long* norm;
    int ytiles, ztiles, ntiles, ntiles_pad;
    int64_t xchunk;
};

constexpr int RZ = 62;  // output lanes per residual tile: lanes 1..62

template <class T, int MODE>
__device__ __forceinline__ T resid_point(T uc, T xm, T xp, T ym, T yp, T zm, T zp, T fc, T hsq, T inv_h2) {
    if (MODE == MGP_MODE_REFERENCE) {
        return fc - (xm + xp + ym + yp + zm + zp - 6.0 * uc) / hsq;  // residual.h:70-72, double literal
    } else {
        const T s = ((xm - uc) + (xp - uc)) + ((ym - uc) + (yp - uc)) + ((zm - uc) + (zp - uc));
        return fma(-inv_h2, s, fc);
    }
}

template <class T, class O, int MODE, int R, bool NORM, bool STORE = true>
__global__ void __launch_bounds__(64) k_residual_wave(ResidArgs<T, O> A) {
    constexpr int NR = R + 2;
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    const int tz = tl % A.ztiles;
    const int ty = (tl / A.ztiles) % A.ytiles;
    const int xc = tl / (A.ztiles * A.ytiles);
    const int64_t y0 = A.by0 + (int64_t)ty * R;
    const int64_t z0 = A.bz0 + (int64_t)tz * RZ;
    const int64_t xa = A.bx0 + (int64_t)xc * A.xchunk;
    const int64_t xb = min(A.bx1, xa + A.xchunk);
    const int64_t s0 = A.u.s0, s1 = A.u.s1;
    // buffer base: (plane, ylo, zlo); every offset below is in range
    const uint32_t plane_bytes = (uint32_t)(((A.yhi - A.ylo) * s1 + (A.zhi - A.zlo) + 1) * (int64_t)sizeof(T));
    const int lane = threadIdx.x;
    const int64_t z = z0 - 1 + lane;
    const bool lane_out = lane >= 1 && lane <= RZ && z < A.bz1;
    const int64_t zc = min(max(z, A.zlo), A.zhi);
    const uint32_t zoff = (uint32_t)((zc - A.zlo) * (int64_t)sizeof(T));
    uint32_t roff[NR];
    bool yout[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) {
        const int64_t y = y0 - 1 + t;
        roff[t] = (uint32_t)((min(max(y, A.ylo), A.yhi) - A.ylo) * s1 * (int64_t)sizeof(T));
        yout[t] = t >= 1 && t <= R && y < A.by1;
    }
    const int64_t corner = A.ylo * s1 + A.zlo;  // element offset of (0, ylo, zlo) from (0, 0, 0)
    auto rsrc = [&](const T* p, int64_t i) {
        return plane_rsrc(p + min(max(i, A.xlo), A.xhi) * s0 + corner, plane_bytes);
    };
    auto load_u = [&](int64_t i, T (&U)[NR]) {
        const __amdgpu_buffer_rsrc_t rs = rsrc(A.u.p, i);
#pragma unroll
        for (int t = 0; t < NR; t++) U[t] = bload<T>(rs, zoff, roff[t]);
    };
    auto load_f = [&](int64_t i, T (&F)[NR]) {
        const __amdgpu_buffer_rsrc_t rs = rsrc(A.f.p, i);
#pragma unroll
        for (int t = 1; t <= R; t++) F[t] = bload<T>(rs, zoff, roff[t]);
    };
    T Um1[NR], U0[NR], Up1[NR], UL[NR], F0[NR], FL[NR];
    load_u(xa - 1, Um1);
    load_u(xa, U0);
    load_u(xa + 1, Up1);
    load_f(xa, F0);
    double vmax = 0.0;
    for (int64_t i = xa; i < xb; i++) {
        load_u(i + 2, UL);
        load_f(i + 1, FL);
        // output plane (element offsets as u's; O may be narrower than T)
        const __amdgpu_buffer_rsrc_t ro = plane_rsrc(
            A.out.p + min(max(i, A.xlo), A.xhi) * s0 + corner,
            (uint32_t)((uint64_t)plane_bytes * sizeof(O) / sizeof(T)));
#pragma unroll
        for (int t = 1; t <= R; t++) {
            const T rv = resid_point<T, MODE>(U0[t], Um1[t], Up1[t], U0[t - 1], U0[t + 1], lane_prev(U0[t]),
                                              lane_next(U0[t]), F0[t], A.hsq, A.inv_h2);
            const bool out = lane_out && yout[t];
            if (STORE)
                bstore((O)rv, ro, out ? (uint32_t)((uint64_t)zoff * sizeof(O) / sizeof(T)) : BUF_OOB,
                       (uint32_t)((uint64_t)roff[t] * sizeof(O) / sizeof(T)));
            if (NORM) vmax = out ? fmax(vmax, fabs((double)rv)) : vmax;
        }
#pragma unroll
        for (int t = 0; t < NR; t++) {
            Um1[t] = U0[t];
            U0[t] = Up1[t];
            Up1[t] = UL[t];
            F0[t] = FL[t];
        }
    }
    if (NORM) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o, 64));
        if (lane == 0 && vmax > 0.0) atomicMax(A.norm, (unsigned long long)__double_as_longlong(vmax));
    }
}

// FP64 update and residual of iterative refinement in one pass
// (mgp_update_residual_f64_f32): u_new = u + (double) e written to a second
// buffer, then r = f - L u_new on the box written as FP32, with max |r|.  The
// wave march of k_residual_wave, each loaded position updated in registers
// first; the Neumann ghosts the boundary refresh would store before the
// residual (include/residual.h:26-34 with include/neumann.h:112-130) are formed
// from u_new in registers: y ghosts from the rows two apart, z ghosts from the
// lanes two apart, x-ghost planes from the plane two apart.  The shape points
// outside the box get u + e from extra blocks after the tiles; u_new's halo is
// not written.
struct UpdArgs {
    View<double> u, un, f, r;
    View<float> e, out;
    const double* g[6];
    double two_h;
    int64_t bx0, bx1, by0, by1, bz0, bz1;
    int64_t ylo, yhi, zlo, zhi, xlo, xhi;  // allocated (halo-inclusive) index ranges
    double hsq, inv_h2;
    unsigned long long* norm;
    int ytiles, ztiles, ntiles, ntiles_pad;
    int64_t xchunk;
};

// shape points outside the box: un = u + (double) e (the order of
// copy_complement)
__device__ void update_complement(const UpdArgs& A, int blk) {
    const int64_t n0 = A.u.n0, n1 = A.u.n1, n2 = A.u.n2;
    const int64_t BX = A.bx1 - A.bx0, BY = A.by1 - A.by0;
    const int64_t xo = A.bx0 + n0 - A.bx1, yo = A.by0 + n1 - A.by1, zo = A.bz0 + n2 - A.bz1;
    const int64_t c1 = xo * n1 * n2, c2 = BX * yo * n2, c3 = BX * BY * zo;
    const int64_t per = (int64_t)blockDim.x * COPY_PER_THREAD;
    const int64_t end = min(c1 + c2 + c3, (int64_t)(blk + 1) * per);
    for (int64_t t = (int64_t)blk * per + threadIdx.x; t < end; t += blockDim.x) {
        int64_t x, y, z;
        if (t < c1) {
            const int64_t a = t / (n1 * n2), r = t % (n1 * n2);
            x = a < A.bx0 ? a : A.bx1 + (a - A.bx0);
            y = r / n2;
            z = r % n2;
        } else if (t < c1 + c2) {
            const int64_t u = t - c1, a = u / (yo * n2), r = u % (yo * n2), c = r / n2;
            x = A.bx0 + a;
            y = c < A.by0 ? c : A.by1 + (c - A.by0);
            z = r % n2;
        } else {
            const int64_t u = t - c1 - c2, a = u / (BY * zo), r = u % (BY * zo), c = r % zo;
            x = A.bx0 + a;
            y = A.by0 + r / zo;
            z = c < A.bz0 ? c : A.bz1 + (c - A.bz0);
        }
        A.un.at(x, y, z) = A.u.at(x, y, z) + (double)A.e.at(x, y, z);
    }
}

template <int MODE, int R>
__global__ void __launch_bounds__(64) k_update_residual_wave(UpdArgs A) {
    typedef double T;
    constexpr int NR = R + 2;
    if ((int)blockIdx.x >= A.ntiles_pad) {
        update_complement(A, (int)blockIdx.x - A.ntiles_pad);
        return;
    }
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    const int tz = tl % A.ztiles;
    const int ty = (tl / A.ztiles) % A.ytiles;
    const int xc = tl / (A.ztiles * A.ytiles);
    const int64_t y0 = A.by0 + (int64_t)ty * R;
    const int64_t z0 = A.bz0 + (int64_t)tz * RZ;
    const int64_t xa = A.bx0 + (int64_t)xc * A.xchunk;
    const int64_t xb = min(A.bx1, xa + A.xchunk);
    const int64_t n0 = A.u.n0, n1 = A.u.n1, n2 = A.u.n2, s0 = A.u.s0, s1 = A.u.s1;
    const double two_h = A.two_h;
    const uint32_t plane_bytes = (uint32_t)(((A.yhi - A.ylo) * s1 + (A.zhi - A.zlo) + 1) * (int64_t)sizeof(T));
    const uint32_t eplane_bytes = plane_bytes / 2;
    const int lane = threadIdx.x;
    const int64_t z = z0 - 1 + lane;
    const bool lane_out = lane >= 1 && lane <= RZ && z < A.bz1;
    const bool z_shape = z >= 0 && z < n2;
    const int64_t zc = min(max(z, A.zlo), A.zhi);
    const int64_t zs = min(max(z, (int64_t)0), n2 - 1);  // clamped into the shape (face index)
    const uint32_t zoff = (uint32_t)((zc - A.zlo) * (int64_t)sizeof(T));
    uint32_t roff[NR];
    bool yout[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) {
        const int64_t y = y0 - 1 + t;
        roff[t] = (uint32_t)((min(max(y, A.ylo), A.yhi) - A.ylo) * s1 * (int64_t)sizeof(T));
        yout[t] = t >= 1 && t <= R && y < A.by1;
    }
    const double* gW = A.g[MGP_WEST];
    const double* gE = A.g[MGP_EAST];
    const double* gS = A.g[MGP_SOUTH];
    const double* gN = A.g[MGP_NORTH];
    const double* gB = A.g[MGP_BOTTOM];
    const double* gT = A.g[MGP_TOP];
    const bool y_ghosts = (y0 - 1 < 0 && gS) || (y0 + R >= n1 && gN);
    const bool z_ghosts = (z0 - 1 < 0 && gB) || (z0 + RZ >= n2 && gT);
    const int64_t corner = A.ylo * s1 + A.zlo;
    // u_new of plane i (clamped into the allocated range) with its in-plane
    // Neumann ghosts
    auto load_un = [&](int64_t i, T (&U)[NR]) {
        const int64_t ic = min(max(i, A.xlo), A.xhi);
        const __amdgpu_buffer_rsrc_t ru = plane_rsrc(A.u.p + ic * s0 + corner, plane_bytes);
        const __amdgpu_buffer_rsrc_t re = plane_rsrc(A.e.p + ic * s0 + corner, eplane_bytes);
#pragma unroll
        for (int t = 0; t < NR; t++) U[t] = bload<T>(ru, zoff, roff[t]) + (double)bload<float>(re, zoff / 2, roff[t] / 2);
    };
    auto ghosts = [&](int64_t i, T (&U)[NR]) {
        if (i < 0 || i >= n0) return;  // x-ghost planes are formed whole below
        if (y_ghosts) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t y = y0 - 1 + t;
                if (t + 2 < NR && y == -1 && gS) {
                    const T v = (T)-1.0 * two_h * gS[i * n2 + zs] + U[t + 2];  // face (nx, 1, nz)
                    U[t] = z_shape ? v : U[t];
                } else if (t >= 2 && y == n1 && gN) {
                    const T v = (T)1.0 * two_h * gN[i * n2 + zs] + U[t - 2];
                    U[t] = z_shape ? v : U[t];
                }
            }
        }
        if (z_ghosts) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t y = y0 - 1 + t;
                if (y < 0 || y >= n1) continue;
                if (gB) {
                    const T v = (T)-1.0 * two_h * gB[i * n1 + y] + lane_next(lane_next(U[t]));  // face (nx, ny, 1)
                    U[t] = z == -1 ? v : U[t];
                }
                if (gT) {
                    const T v = (T)1.0 * two_h * gT[i * n1 + y] + lane_prev(lane_prev(U[t]));
                    U[t] = z == n2 ? v : U[t];
                }
            }
        }
    };
    auto load_f = [&](int64_t i, T (&F)[NR]) {
        const __amdgpu_buffer_rsrc_t rs = plane_rsrc(A.f.p + min(max(i, A.xlo), A.xhi) * s0 + corner, plane_bytes);
#pragma unroll
        for (int t = 1; t <= R; t++) F[t] = bload<T>(rs, zoff, roff[t]);
    };
    T Um1[NR], U0[NR], Up1[NR], UL[NR], F0[NR], FL[NR];
    load_un(xa - 1, Um1);
    load_un(xa, U0);
    load_un(xa + 1, Up1);
    load_f(xa, F0);
    ghosts(xa - 1, Um1);
    ghosts(xa, U0);
    double vmax = 0.0;
    for (int64_t i = xa; i < xb; i++) {
        load_un(i + 2, UL);
        load_f(i + 1, FL);
        ghosts(i + 1, Up1);
        // x-ghost planes from the plane two apart (face layout (1, ny, nz))
        if (i == 0 && gW) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t yc = min(max(y0 - 1 + t, (int64_t)0), n1 - 1);
                Um1[t] = (T)-1.0 * two_h * gW[yc * n2 + zs] + Up1[t];
            }
        }
        if (i == n0 - 1 && gE) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t yc = min(max(y0 - 1 + t, (int64_t)0), n1 - 1);
                Up1[t] = (T)1.0 * two_h * gE[yc * n2 + zs] + Um1[t];
            }
        }
        const int64_t ic = min(max(i, A.xlo), A.xhi);
        const __amdgpu_buffer_rsrc_t rn = plane_rsrc(A.un.p + ic * s0 + corner, plane_bytes);
        const __amdgpu_buffer_rsrc_t ro = plane_rsrc(A.out.p + ic * s0 + corner, eplane_bytes);
#pragma unroll
        for (int t = 1; t <= R; t++) {
            const bool out = lane_out && yout[t];
            bstore(U0[t], rn, out ? zoff : BUF_OOB, roff[t]);
            const T rv = resid_point<T, MODE>(U0[t], Um1[t], Up1[t], U0[t - 1], U0[t + 1], lane_prev(U0[t]),
                                              lane_next(U0[t]), F0[t], A.hsq, A.inv_h2);
            bstore((float)rv, ro, out ? zoff / 2 : BUF_OOB, roff[t] / 2);
            vmax = out ? fmax(vmax, fabs(rv)) : vmax;
        }
#pragma unroll
        for (int t = 0; t < NR; t++) {
            Um1[t] = U0[t];
            U0[t] = Up1[t];
            Up1[t] = UL[t];
            F0[t] = FL[t];
        }
    }
    if (A.norm) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o, 64));
        if (lane == 0 && vmax > 0.0) atomicMax(A.norm, (unsigned long long)__double_as_longlong(vmax));
    }
}

// The iterate of mixed-precision refinement held as hi + lo (FP64 hi, FP32
// lo; mgp_update_residual_hilo_f64_f32, mgp_residual_hilo_f64).  UPD: the
// update hi + lo + e, renormalised by an error-free sum, goes to a second pair
// of buffers (hn, ln), then r = f - L(hi + lo) of the updated iterate is
// written as FP32 into `out` (the next right-hand side); !UPD: r of the
// stored iterate written as FP64 into `out`.  The march of
// k_update_residual_wave; Neumann ghosts of hi + lo formed in registers with
// the rounding of p + hi(mirror) carried into lo.  Consistent mode only.
struct HiloArgs {
    View<double> u, un, f, r;
    View<float> lo, ln, e;
    double* out64;
    float* out32;
    const double* g[6];
    double two_h;
    int64_t bx0, bx1, by0, by1, bz0, bz1;
    int64_t ylo, yhi, zlo, zhi, xlo, xhi;
    double inv_h2;
    unsigned long long* norm;
    int ytiles, ztiles, ntiles, ntiles_pad;
    int64_t xchunk;
};

// shape points outside the box: (hn, ln) = hi + lo + e (the order of
// copy_complement)
__device__ void hilo_complement(const HiloArgs& A, int blk) {
    const int64_t n0 = A.u.n0, n1 = A.u.n1, n2 = A.u.n2;
    const int64_t BX = A.bx1 - A.bx0, BY = A.by1 - A.by0;
    const int64_t xo = A.bx0 + n0 - A.bx1, yo = A.by0 + n1 - A.by1, zo = A.bz0 + n2 - A.bz1;
    const int64_t c1 = xo * n1 * n2, c2 = BX * yo * n2, c3 = BX * BY * zo;
    const int64_t per = (int64_t)blockDim.x * COPY_PER_THREAD;
    const int64_t end = min(c1 + c2 + c3, (int64_t)(blk + 1) * per);
    for (int64_t t = (int64_t)blk * per + threadIdx.x; t < end; t += blockDim.x) {
        int64_t x, y, z;
        if (t < c1) {
            const int64_t a = t / (n1 * n2), r = t % (n1 * n2);
            x = a < A.bx0 ? a : A.bx1 + (a - A.bx0);
            y = r / n2;
            z = r % n2;
        } else if (t < c1 + c2) {
            const int64_t u = t - c1, a = u / (yo * n2), r = u % (yo * n2), c = r / n2;
            x = A.bx0 + a;
            y = c < A.by0 ? c : A.by1 + (c - A.by0);
            z = r % n2;
        } else {
            const int64_t u = t - c1 - c2, a = u / (BY * zo), r = u % (BY * zo), c = r % zo;
            x = A.bx0 + a;
            y = A.by0 + r / zo;
            z = c < A.bz0 ? c : A.bz1 + (c - A.bz0);
        }
        double hi = A.u.at(x, y, z);
        float lo = A.lo.at(x, y, z);
        hilo_update(hi, lo, A.e.at(x, y, z));
        A.un.at(x, y, z) = hi;
        A.ln.at(x, y, z) = lo;
    }
}

template <int R, bool UPD>
__global__ void __launch_bounds__(64) k_hilo_residual_wave(HiloArgs A) {
    constexpr int NR = R + 2;
    if ((int)blockIdx.x >= A.ntiles_pad) {
        if (UPD) hilo_complement(A, (int)blockIdx.x - A.ntiles_pad);
        return;
    }
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    const int tz = tl % A.ztiles;
    const int ty = (tl / A.ztiles) % A.ytiles;
    const int xc = tl / (A.ztiles * A.ytiles);
    const int64_t y0 = A.by0 + (int64_t)ty * R;
    const int64_t z0 = A.bz0 + (int64_t)tz * RZ;
    const int64_t xa = A.bx0 + (int64_t)xc * A.xchunk;
    const int64_t xb = min(A.bx1, xa + A.xchunk);
    const int64_t n0 = A.u.n0, n1 = A.u.n1, n2 = A.u.n2, s0 = A.u.s0, s1 = A.u.s1;
    const double two_h = A.two_h;
    const uint32_t plane_bytes = (uint32_t)(((A.yhi - A.ylo) * s1 + (A.zhi - A.zlo) + 1) * (int64_t)sizeof(double));
    const uint32_t fplane_bytes = plane_bytes / 2;
    const int lane = threadIdx.x;
    const int64_t z = z0 - 1 + lane;
    const bool lane_out = lane >= 1 && lane <= RZ && z < A.bz1;
    const bool z_shape = z >= 0 && z < n2;
    const int64_t zc = min(max(z, A.zlo), A.zhi);
    const int64_t zs = min(max(z, (int64_t)0), n2 - 1);  // clamped into the shape (face index)
    const uint32_t zoff = (uint32_t)((zc - A.zlo) * (int64_t)sizeof(double));
    uint32_t roff[NR];
    bool yout[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) {
        const int64_t y = y0 - 1 + t;
        roff[t] = (uint32_t)((min(max(y, A.ylo), A.yhi) - A.ylo) * s1 * (int64_t)sizeof(double));
        yout[t] = t >= 1 && t <= R && y < A.by1;
    }
    const double* gW = A.g[MGP_WEST];
    const double* gE = A.g[MGP_EAST];
    const double* gS = A.g[MGP_SOUTH];
    const double* gN = A.g[MGP_NORTH];
    const double* gB = A.g[MGP_BOTTOM];
    const double* gT = A.g[MGP_TOP];
    const bool y_ghosts = (y0 - 1 < 0 && gS) || (y0 + R >= n1 && gN);
    const bool z_ghosts = (z0 - 1 < 0 && gB) || (z0 + RZ >= n2 && gT);
    const int64_t corner = A.ylo * s1 + A.zlo;
    // hi + lo (+ e) of plane i, clamped into the allocated range
    auto load_x = [&](int64_t i, double (&H)[NR], float (&L)[NR]) {
        const int64_t ic = min(max(i, A.xlo), A.xhi);
        const __amdgpu_buffer_rsrc_t rh = plane_rsrc(A.u.p + ic * s0 + corner, plane_bytes);
        const __amdgpu_buffer_rsrc_t rl = plane_rsrc(A.lo.p + ic * s0 + corner, fplane_bytes);
#pragma unroll
        for (int t = 0; t < NR; t++) {
            H[t] = bload<double>(rh, zoff, roff[t]);
            L[t] = bload<float>(rl, zoff / 2, roff[t] / 2);
        }
        if (UPD) {
            const __amdgpu_buffer_rsrc_t re = plane_rsrc(A.e.p + ic * s0 + corner, fplane_bytes);
#pragma unroll
            for (int t = 0; t < NR; t++) hilo_update(H[t], L[t], bload<float>(re, zoff / 2, roff[t] / 2));
        }
    };
    auto ghosts = [&](int64_t i, double (&H)[NR], float (&L)[NR]) {
        if (i < 0 || i >= n0) return;  // x-ghost planes are formed whole below
        if (y_ghosts) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t y = y0 - 1 + t;
                double hg;
                float lg;
                if (t + 2 < NR && y == -1 && gS) {
                    hilo_ghost((double)-1.0 * two_h * gS[i * n2 + zs], H[t + 2], L[t + 2], hg, lg);  // (nx, 1, nz)
                    H[t] = z_shape ? hg : H[t];
                    L[t] = z_shape ? lg : L[t];
                } else if (t >= 2 && y == n1 && gN) {
                    hilo_ghost((double)1.0 * two_h * gN[i * n2 + zs], H[t - 2], L[t - 2], hg, lg);
                    H[t] = z_shape ? hg : H[t];
                    L[t] = z_shape ? lg : L[t];
                }
            }
        }
        if (z_ghosts) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t y = y0 - 1 + t;
                if (y < 0 || y >= n1) continue;
                double hg;
                float lg;
                if (gB) {
                    hilo_ghost((double)-1.0 * two_h * gB[i * n1 + y], lane_next(lane_next(H[t])),
                               lane_next(lane_next(L[t])), hg, lg);  // face (nx, ny, 1)
                    H[t] = z == -1 ? hg : H[t];
                    L[t] = z == -1 ? lg : L[t];
                }
                if (gT) {
                    hilo_ghost((double)1.0 * two_h * gT[i * n1 + y], lane_prev(lane_prev(H[t])),
                               lane_prev(lane_prev(L[t])), hg, lg);
                    H[t] = z == n2 ? hg : H[t];
                    L[t] = z == n2 ? lg : L[t];
                }
            }
        }
    };
    auto load_f = [&](int64_t i, double (&F)[NR]) {
        const __amdgpu_buffer_rsrc_t rs = plane_rsrc(A.f.p + min(max(i, A.xlo), A.xhi) * s0 + corner, plane_bytes);
#pragma unroll
        for (int t = 1; t <= R; t++) F[t] = bload<double>(rs, zoff, roff[t]);
    };
    double Hm1[NR], H0[NR], Hp1[NR], HL[NR], F0[NR], FL[NR];
    float Lm1[NR], L0[NR], Lp1[NR], LL[NR];
    load_x(xa - 1, Hm1, Lm1);
    load_x(xa, H0, L0);
    load_x(xa + 1, Hp1, Lp1);
    load_f(xa, F0);
    ghosts(xa - 1, Hm1, Lm1);
    ghosts(xa, H0, L0);
    double vmax = 0.0;
    for (int64_t i = xa; i < xb; i++) {
        load_x(i + 2, HL, LL);
        load_f(i + 1, FL);
        ghosts(i + 1, Hp1, Lp1);
        // x-ghost planes from the plane two apart (face layout (1, ny, nz))
        if (i == 0 && gW) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t yc = min(max(y0 - 1 + t, (int64_t)0), n1 - 1);
                hilo_ghost((double)-1.0 * two_h * gW[yc * n2 + zs], Hp1[t], Lp1[t], Hm1[t], Lm1[t]);
            }
        }
        if (i == n0 - 1 && gE) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
                const int64_t yc = min(max(y0 - 1 + t, (int64_t)0), n1 - 1);
                hilo_ghost((double)1.0 * two_h * gE[yc * n2 + zs], Hm1[t], Lm1[t], Hp1[t], Lp1[t]);
            }
        }
        const int64_t ic = min(max(i, A.xlo), A.xhi);
        __amdgpu_buffer_rsrc_t rn, rln, ro;
        if (UPD) {
            rn = plane_rsrc(A.un.p + ic * s0 + corner, plane_bytes);
            rln = plane_rsrc(A.ln.p + ic * s0 + corner, fplane_bytes);
            ro = plane_rsrc(A.out32 + ic * s0 + corner, fplane_bytes);
        } else {
            ro = plane_rsrc(A.out64 + ic * s0 + corner, plane_bytes);
        }
#pragma unroll
        for (int t = 1; t <= R; t++) {
            const bool out = lane_out && yout[t];
            const double rv = hilo_residual_point(H0[t], Hm1[t], Hp1[t], H0[t - 1], H0[t + 1], lane_prev(H0[t]),
                                                  lane_next(H0[t]), L0[t], Lm1[t], Lp1[t], L0[t - 1], L0[t + 1],
                                                  lane_prev(L0[t]), lane_next(L0[t]), F0[t], A.inv_h2);
            if (UPD) {
                bstore(H0[t], rn, out ? zoff : BUF_OOB, roff[t]);
                bstore(L0[t], rln, out ? zoff / 2 : BUF_OOB, roff[t] / 2);
                bstore((float)rv, ro, out ? zoff / 2 : BUF_OOB, roff[t] / 2);
            } else {
                bstore(rv, ro, out ? zoff : BUF_OOB, roff[t]);
            }
            vmax = out ? fmax(vmax, fabs(rv)) : vmax;
        }
#pragma unroll
        for (int t = 0; t < NR; t++) {
            Hm1[t] = H0[t];
            H0[t] = Hp1[t];
            Hp1[t] = HL[t];
            Lm1[t] = L0[t];
            L0[t] = Lp1[t];
            Lp1[t] = LL[t];
            F0[t] = FL[t];
        }
    }
    if (A.norm) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o, 64));
        if (lane == 0 && vmax > 0.0) atomicMax(A.norm, (unsigned long long)__double_as_longlong(vmax));
    }
}

// max |r| over a region (the part of the shape outside the box), into *norm
template <class T>
__global__ void k_norm_region(View<T> r, int64_t x0, int64_t x1, int64_t y0, int64_t y1, int64_t z0, int64_t z1,
                              unsigned long long* norm) {
    const int64_t ny = y1 - y0, nz = z1 - z0;
    const int64_t n = (x1 - x0) * ny * nz;
    double v = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = x0 + t / (ny * nz), j = y0 + (t / nz) % ny, k = z0 + t % nz;
        v = fmax(v, fabs((double)r.at(i, j, k)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(norm, (unsigned long long)__double_as_longlong(v));
}


// ---------------------------------------------------------------------------
// Trilinear prolongation added to the fine solution, u += P(coarse), over the
// fine shape (include/trilinearinterpolation.h:50-133 composed as the four
// passes compose it, then include/devicearray.h:175-194).  One wave per fine
// (y, z) tile of R rows x 64 columns marching along x: the fine plane is read,
// updated and written; the coarse values it needs (coarse rows b and b+1,
// columns c and c+1, planes a and a+1) are loaded into registers once per
// coarse plane and shared by the two fine planes that use them.
template <class T>
struct ProlongArgs {
    View<T> c, u;  // coarse, fine
    int ytiles, ztiles, ntiles, ntiles_pad;
    int64_t xchunk;  // even
};

template <class T, int R>
__global__ void __launch_bounds__(64) k_prolong_add_wave(ProlongArgs<T> A) {
    static_assert(R % 2 == 0, "fine tiles start on even rows");
    constexpr int CR = R / 2 + 1;  // coarse rows a tile reads
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    const int tz = tl % A.ztiles;
    const int ty = (tl / A.ztiles) % A.ytiles;
    const int xc = tl / (A.ztiles * A.ytiles);
    const int64_t n0 = A.u.n0, n1 = A.u.n1, n2 = A.u.n2;
    const int64_t c0 = A.c.n0, c1 = A.c.n1, c2 = A.c.n2;
    const int64_t y0 = (int64_t)ty * R, z0 = (int64_t)tz * 64;
    const int64_t xa = (int64_t)xc * A.xchunk, xb = min(n0, xa + A.xchunk);
    const int lane = threadIdx.x;
    const int64_t k = z0 + lane;
    const bool lane_in = k < n2;
    const int64_t kc = min(k, n2 - 1);
    const int pk = (int)(kc & 1);
    const uint32_t zoff = (uint32_t)(kc * (int64_t)sizeof(T));
    const uint32_t zoff_st = lane_in ? zoff : BUF_OOB;
    // coarse columns c = k>>1 and c+1 (clamped; c+1 is only used for odd k)
    const uint32_t czo = (uint32_t)(min(kc >> 1, c2 - 1) * (int64_t)sizeof(T));
    const uint32_t czo1 = (uint32_t)(min((kc >> 1) + 1, c2 - 1) * (int64_t)sizeof(T));
    const uint32_t fplane = (uint32_t)(((n1 - 1) * A.u.s1 + n2) * (int64_t)sizeof(T));
    const uint32_t cplane = (uint32_t)(((c1 - 1) * A.c.s1 + c2) * (int64_t)sizeof(T));
    uint32_t froff[R], croff[CR];
    bool row_in[R];
#pragma unroll
    for (int t = 0; t < R; t++) {
        const int64_t j = y0 + t;
        row_in[t] = j < n1;
        froff[t] = (uint32_t)(min(j, n1 - 1) * A.u.s1 * (int64_t)sizeof(T));
    }
#pragma unroll
    for (int t = 0; t < CR; t++) croff[t] = (uint32_t)(min((y0 >> 1) + t, c1 - 1) * A.c.s1 * (int64_t)sizeof(T));
    // coarse plane a: values at (row b, column c) and (row b, column c+1)
    auto load_coarse = [&](int64_t a, T (&V)[CR], T (&V1)[CR]) {
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(A.c.p + min(a, c0 - 1) * A.c.s0, cplane);
#pragma unroll
        for (int t = 0; t < CR; t++) {
            V[t] = bload<T>(r, czo, croff[t]);
            V1[t] = bload<T>(r, czo1, croff[t]);
        }
    };
    auto half_sum = [](T a, T b) -> T { return 0.5 * (a + b); };  // trilinearinterpolation.h: 0.5 * (l + r)
    // xy value at coarse column set (V: plane a, W: plane a+1) for fine row t
    // of parity pj, fine plane parity pi (the reference's x-then-y order)
    auto tri_xy = [&](const T (&V)[CR], const T (&W)[CR], int t, int pi) -> T {
        const int b = t >> 1, pj = t & 1;  // y0 is even
        if (!pi && !pj) return V[b];
        if (pi && !pj) return half_sum(V[b], W[b]);
        if (!pi && pj) return half_sum(V[b], V[b + 1]);
        return half_sum(half_sum(V[b], W[b]), half_sum(V[b + 1], W[b + 1]));
    };
    T Ca[CR], Ca1[CR], Cb[CR], Cb1[CR], U[R], UL[R];
    load_coarse(xa >> 1, Ca, Ca1);
    load_coarse((xa >> 1) + 1, Cb, Cb1);
    auto load_fine = [&](int64_t i, T (&D)[R]) {
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(A.u.p + min(i, n0 - 1) * A.u.s0, fplane);
#pragma unroll
        for (int t = 0; t < R; t++) D[t] = bload<T>(r, zoff, froff[t]);
    };
    load_fine(xa, U);
    for (int64_t i = xa; i < xb; i++) {
        load_fine(i + 1, UL);
        const int pi = (int)(i & 1);
        const __amdgpu_buffer_rsrc_t ro = plane_rsrc(A.u.p + i * A.u.s0, fplane);
#pragma unroll
        for (int t = 0; t < R; t++) {
            const T lo = tri_xy(Ca, Cb, t, pi);
            T v = lo;
            if (pk) v = half_sum(lo, tri_xy(Ca1, Cb1, t, pi));
            bstore((T)(U[t] + v), ro, row_in[t] ? zoff_st : BUF_OOB, froff[t]);
        }
#pragma unroll
        for (int t = 0; t < R; t++) U[t] = UL[t];
        if (pi) {  // the next fine plane is even: coarse plane a+1 becomes a
#pragma unroll
            for (int t = 0; t < CR; t++) Ca[t] = Cb[t], Ca1[t] = Cb1[t];
            load_coarse(((i + 1) >> 1) + 1, Cb, Cb1);
        }
    }
}


// ---------------------------------------------------------------------------
// Residual and full-weighting restriction in one pass (consistent mode):
// coarse f = FW(r) over the coarse shape with r = f - L u computed on the fine
// box in registers and never stored (the fine r is dead after the restriction
// when the prolongation is fused), r outside the box read as stored, and the
// neighbour indices clamped to the shape (include/fullweighting.h:59-121,
// include/residual.h:58-79).  One wave per coarse tile of RJ rows x 30
// columns marching along x: fine u rows 2J0-2 .. 2J0+2RJ (NU), r rows
// 2J0-1 .. 2J0+2RJ-1 (NRr), fine lanes z = 2K0-2+l; coarse column K sits on
// lane 2(K-K0)+2.  Coarse plane I is emitted after r of fine plane 2I+1.
template <class T>
struct RFArgs {
    View<T> u, f, r, c;  // fine u, f, r; coarse f
    int64_t bx0, bx1, by0, by1, bz0, bz1;
    int64_t ylo, yhi, zlo, zhi, xlo, xhi;  // allocated index ranges of the fine arrays
    T hsq, inv_h2;
    int ytiles, ztiles, ntiles, ntiles_pad;
    int64_t ichunk;  // coarse planes per wave
};

constexpr int FWK = 30;  // coarse columns per wave

// D: fine planes of loads in flight ahead of the residual (1 or 2)
template <class T, int RJ, int D, bool EDGE>
__device__ __forceinline__ void resid_fw_march(const RFArgs<T>& A, int tl) {
    constexpr int NU = 2 * RJ + 3, NRr = 2 * RJ + 1;
    const int tz = tl % A.ztiles;
    const int ty = (tl / A.ztiles) % A.ytiles;
    const int xc = tl / (A.ztiles * A.ytiles);
    const int64_t n0 = A.u.n0, n1 = A.u.n1, n2 = A.u.n2;
    const int64_t c0 = A.c.n0, c1 = A.c.n1, c2 = A.c.n2;
    const int64_t J0 = (int64_t)ty * RJ, K0 = (int64_t)tz * FWK;
    const int64_t Ia = (int64_t)xc * A.ichunk, Ib = min(c0, Ia + A.ichunk);
    const int64_t s0 = A.u.s0, s1 = A.u.s1;
    const uint32_t plane_bytes = (uint32_t)(((A.yhi - A.ylo) * s1 + (A.zhi - A.zlo) + 1) * (int64_t)sizeof(T));
    const int64_t corner = A.ylo * s1 + A.zlo;
    const int lane = threadIdx.x;
    const int64_t z = 2 * K0 - 2 + lane;
    const int64_t zc = min(max(z, A.zlo), A.zhi);
    const uint32_t zoff = (uint32_t)((zc - A.zlo) * (int64_t)sizeof(T));
    const bool z_box = z >= A.bz0 && z < A.bz1 && lane >= 1 && lane <= 62;
    const bool z_shape = z >= 0 && z < n2;
    uint32_t roff[NU];
    bool y_box[NU], y_shape[NU];
#pragma unroll
    for (int t = 0; t < NU; t++) {
        const int64_t y = 2 * J0 - 2 + t;
        roff[t] = (uint32_t)((min(max(y, A.ylo), A.yhi) - A.ylo) * s1 * (int64_t)sizeof(T));
        y_box[t] = y >= A.by0 && y < A.by1;
        y_shape[t] = y >= 0 && y < n1;
    }
    auto rsrc = [&](const T* p, int64_t i) {
        return plane_rsrc(p + min(max(i, A.xlo), A.xhi) * s0 + corner, plane_bytes);
    };
    auto load_u = [&](int64_t i, T (&U)[NU]) {
        const __amdgpu_buffer_rsrc_t rs = rsrc(A.u.p, i);
#pragma unroll
        for (int t = 0; t < NU; t++) U[t] = bload<T>(rs, zoff, roff[t]);
    };
    auto load_f = [&](int64_t i, T (&F)[NU]) {
        const __amdgpu_buffer_rsrc_t rs = rsrc(A.f.p, i);
#pragma unroll
        for (int t = 1; t <= NRr; t++) F[t] = bload<T>(rs, zoff, roff[t]);
    };
    // r of fine plane i at rows t = 1..NRr (r row t-1): computed on the box,
    // as stored elsewhere in the shape; outside the shape the clamped
    // neighbour's value (rows and lanes; planes are handled at emission)
    auto r_plane = [&](int64_t i, const T (&Um)[NU], const T (&U0)[NU], const T (&Up)[NU], const T (&F0)[NU],
                       T (&Rr)[NRr]) {
        const bool x_box = i >= A.bx0 && i < A.bx1;
        __amdgpu_buffer_rsrc_t rs;
        if (EDGE) rs = rsrc(A.r.p, i);
#pragma unroll
        for (int t = 1; t <= NRr; t++) {
            const T v = resid_point<T, MGP_MODE_CONSISTENT>(U0[t], Um[t], Up[t], U0[t - 1], U0[t + 1],
                                                            lane_prev(U0[t]), lane_next(U0[t]), F0[t], A.hsq,
                                                            A.inv_h2);
            if (!EDGE) {
                Rr[t - 1] = v;
            } else {
                const bool in = x_box && y_box[t] && z_box;
                const T st = bload<T>(rs, zoff, roff[t]);  // stored r (in range: clamped offsets)
                Rr[t - 1] = in ? v : st;
            }
        }
        if (EDGE) {
            // rows outside the shape take the clamped row's value (y = -1 -> 0, y = n1 -> n1-1)
#pragma unroll
            for (int t = 1; t <= NRr; t++) {
                const int64_t y = 2 * J0 - 2 + t;
                if (y < 0 && t < NRr) Rr[t - 1] = Rr[t];
                if (y >= n1 && t > 1) Rr[t - 1] = Rr[t - 2];
            }
            // lanes outside the shape: z = -1 -> lane+1, z = n2 -> lane-1
#pragma unroll
            for (int t = 0; t < NRr; t++) {
                const T nx = lane_next(Rr[t]), pv = lane_prev(Rr[t]);
                Rr[t] = z == -1 ? nx : (z == n2 ? pv : Rr[t]);
            }
        }
        (void)z_shape;
        (void)y_shape;
    };
    T Um[NU], U0[NU], Up[NU], Up2[NU], UL[NU], F0[NU], F1[NU], FL[NU], R2[NRr], R1[NRr], R0[NRr];
    const int64_t i0 = 2 * Ia - 1;  // first fine plane whose r is needed
    load_u(i0 - 1, Um);
    load_u(i0, U0);
    load_u(i0 + 1, Up);
    load_f(i0, F0);
    if (D == 2) {
        load_u(i0 + 2, Up2);
        load_f(i0 + 1, F1);
    }
#pragma unroll
    for (int t = 0; t < NRr; t++) R1[t] = R0[t] = (T)0;
    const uint32_t cplane = (uint32_t)(((c1 - 1) * A.c.s1 + c2) * (int64_t)sizeof(T));
    const int64_t K = K0 + (lane - 2) / 2;
    const bool k_out = lane >= 2 && lane <= 60 && (lane & 1) == 0 && K < c2;
    const uint32_t ckoff = (uint32_t)(min(max(K, (int64_t)0), c2 - 1) * (int64_t)sizeof(T));
    for (int64_t i = i0; i <= 2 * Ib - 1; i++) {
        load_u(i + 1 + D, UL);
        load_f(i + D, FL);
        // r(i): planes outside the shape are never emitted directly (clamped below)
#pragma unroll
        for (int t = 0; t < NRr; t++) R2[t] = R1[t], R1[t] = R0[t];
        r_plane(i, Um, U0, Up, F0, R0);
        if ((i & 1) && i >= 2 * Ia + 1) {
            // coarse plane I = (i-1)/2 from r(2I-1) = R2, r(2I) = R1, r(2I+1) = R0,
            // with 2I-1 < 0 and 2I+1 >= n0 clamped to 2I
            const int64_t I = (i - 1) / 2;
            const bool lo_clamp = 2 * I - 1 < 0, hi_clamp = 2 * I + 1 >= n0;
            const __amdgpu_buffer_rsrc_t co = plane_rsrc(A.c.p + I * A.c.s0, cplane);
#pragma unroll
            for (int jr = 0; jr < RJ; jr++) {
                const int64_t J = J0 + jr;
                const int sm = 2 * jr, s0r = 2 * jr + 1, sp = 2 * jr + 2;
                const T* Pm = lo_clamp ? R1 : R2;
                const T* Pp = hi_clamp ? R1 : R0;
                auto A3 = [&](const T* P, int srow, int dz) -> T {
                    const T v = P[srow];
                    return dz < 0 ? lane_prev(v) : (dz > 0 ? lane_next(v) : v);
                };
#define A_(P, srow, dz) A3(P, srow, dz)
                T tmp = 1.0 * A_(Pm, sm, -1) + 2.0 * A_(Pm, s0r, -1) + 1.0 * A_(Pm, sp, -1) + 2.0 * A_(R1, sm, -1) +
                        4.0 * A_(R1, s0r, -1) + 2.0 * A_(R1, sp, -1) + 1.0 * A_(Pp, sm, -1) + 2.0 * A_(Pp, s0r, -1) +
                        1.0 * A_(Pp, sp, -1);
                tmp += 2.0 * A_(Pm, sm, 0) + 4.0 * A_(Pm, s0r, 0) + 2.0 * A_(Pm, sp, 0) + 4.0 * A_(R1, sm, 0) +
                       8.0 * A_(R1, s0r, 0) + 4.0 * A_(R1, sp, 0) + 2.0 * A_(Pp, sm, 0) + 4.0 * A_(Pp, s0r, 0) +
                       2.0 * A_(Pp, sp, 0);
                tmp += 1.0 * A_(Pm, sm, 1) + 2.0 * A_(Pm, s0r, 1) + 1.0 * A_(Pm, sp, 1) + 2.0 * A_(R1, sm, 1) +
                       4.0 * A_(R1, s0r, 1) + 2.0 * A_(R1, sp, 1) + 1.0 * A_(Pp, sm, 1) + 2.0 * A_(Pp, s0r, 1) +
                       1.0 * A_(Pp, sp, 1);
#undef A_
                const T out = tmp / 64.0;
                const uint32_t jo = (uint32_t)(min(J, c1 - 1) * A.c.s1 * (int64_t)sizeof(T));
                bstore(out, co, (k_out && J < c1) ? ckoff : BUF_OOB, jo);
            }
        }
#pragma unroll
        for (int t = 0; t < NU; t++) {
            Um[t] = U0[t];
            U0[t] = Up[t];
            if (D == 2) {
                Up[t] = Up2[t];
                Up2[t] = UL[t];
                F0[t] = F1[t];
                F1[t] = FL[t];
            } else {
                Up[t] = UL[t];
                F0[t] = FL[t];
            }
        }
    }
}

template <class T, int RJ, int D>
__global__ void __launch_bounds__(64) k_resid_fw_wave(RFArgs<T> A) {
    const int per_xcd = A.ntiles_pad / 8;
    const int tl = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (tl >= A.ntiles) return;
    const int tz = tl % A.ztiles;
    const int ty = (tl / A.ztiles) % A.ytiles;
    const int xc = tl / (A.ztiles * A.ytiles);
    // every fine position the tile touches inside the box and the shape
    const int64_t J0 = (int64_t)ty * RJ, K0 = (int64_t)tz * FWK;
    const int64_t Ia = (int64_t)xc * A.ichunk, Ib = min(A.c.n0, Ia + A.ichunk);
    const bool inner = 2 * J0 - 1 >= A.by0 && 2 * (J0 + RJ) - 1 < A.by1 && 2 * K0 - 1 >= A.bz0 &&
                       2 * K0 + 2 * FWK - 1 < A.bz1 && 2 * Ia - 1 >= A.bx0 && 2 * Ib - 1 < A.bx1 &&
                       J0 + RJ <= A.c.n1 && K0 + FWK <= A.c.n2;
    if (inner) resid_fw_march<T, RJ, D, false>(A, tl);
    else resid_fw_march<T, RJ, D, true>(A, tl);
}

}  // namespace

template <class T>
int march_sweep(int mode, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* faces,
                const int64_t b[6], double omega, int64_t xchunk, void* stream) {
    MarchArgs<T> A;
    A.ui = view_of<T>(*ui);
    A.uo = view_of<T>(*uo);
    A.f = view_of<T>(*f);
    for (int q = 0; q < 6; q++) A.g[q] = faces->type[q] == MGP_NEUMANN ? static_cast<const T*>(faces->g[q]) : nullptr;
    A.two_h = (T)(2.0 * faces->h);  // include/neumann.h:108
    A.bx0 = b[0], A.bx1 = b[1], A.by0 = b[2], A.by1 = b[3], A.bz0 = b[4], A.bz1 = b[5];
    A.lin0 = ui->halo[MGP_WEST] * ui->stride[0] + ui->halo[MGP_SOUTH] * ui->stride[1] + ui->halo[MGP_BOTTOM];
    const T om = (T)omega;  // include/gaussseidel.h:50-52
    A.one_omega = (T)(1.0 - om);
    A.omega_sixth = (T)(om / 6.0);
    A.hsq = (T)(faces->h * faces->h);
    A.ytiles = (int)((b[3] - b[2] + TY - 1) / TY);
    A.ztiles = (int)((b[5] - b[4] + TZ - 1) / TZ);
    const int64_t nx = b[1] - b[0];
    if (xchunk <= 0) xchunk = nx;
    const int xchunks = (int)((nx + xchunk - 1) / xchunk);
    A.xchunk = xchunk;
    A.ntiles = A.ytiles * A.ztiles * xchunks;
    A.ntiles_pad = (A.ntiles + 7) / 8 * 8;
    const hipStream_t st = (hipStream_t)stream;
    const bool ref = mode == MGP_MODE_REFERENCE;
    // wave tiles: planes addressed through buffer descriptors with a 32-bit
    // range (a plane must stay below the out-of-range lane offset)
    const int64_t plane_bytes = ((A.ui.n1 - 1) * A.ui.s1 + A.ui.n2) * (int64_t)sizeof(T);
    if (g_sweep_kernel == 2 && A.ui.n0 >= 3 && plane_bytes < (int64_t)BUF_OOB) {
        // register rows per wave tile: measured best 12 (FP64) and 16 (FP32)
        // at 1025 x 1025 x 513
        const int rows = g_sweep_rows == 8 || g_sweep_rows == 12 || g_sweep_rows == 16 ? (int)g_sweep_rows
                                                                                        : (sizeof(T) == 8 ? 12 : 16);
        A.ytiles = (int)((b[3] - b[2] + rows - 3) / (rows - 2));
        A.ztiles = (int)((b[5] - b[4] + WZ - 1) / WZ);
        const int64_t tiles = (int64_t)A.ytiles * A.ztiles;
        int64_t chunks = std::max<int64_t>(1, g_sweep_target_wgs / std::max<int64_t>(tiles, 1));
        chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, nx / std::max<int64_t>(min_chunk_for(box_points(b)), 1)));
        A.xchunk = (nx + chunks - 1) / chunks;
        A.ntiles = (int)(tiles * ((nx + A.xchunk - 1) / A.xchunk));
        A.ntiles_pad = (A.ntiles + 7) / 8 * 8;
        const dim3 g(A.ntiles_pad + complement_blocks(A.ui, b, 64));
#define MGP_WAVE_LAUNCH(RR, DD)                                                                          \
    if (ref) hipLaunchKernelGGL((k_rbgs_wave<T, MGP_MODE_REFERENCE, RR, DD>), g, dim3(64), 0, st, A); \
    else hipLaunchKernelGGL((k_rbgs_wave<T, MGP_MODE_CONSISTENT, RR, DD>), g, dim3(64), 0, st, A);
        const bool d2 = g_sweep_depth == 2;
        if (rows == 8) { if (d2) { MGP_WAVE_LAUNCH(8, 2) } else { MGP_WAVE_LAUNCH(8, 1) } }
        else if (rows == 12) { if (d2) { MGP_WAVE_LAUNCH(12, 2) } else { MGP_WAVE_LAUNCH(12, 1) } }
        else { if (d2) { MGP_WAVE_LAUNCH(16, 2) } else { MGP_WAVE_LAUNCH(16, 1) } }
#undef MGP_WAVE_LAUNCH
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? MGP_OK : (int)e;
    }
    const int threads = g_sweep_waves == 4 ? 256 : g_sweep_waves == 16 ? 1024 : 512;
    const dim3 g(A.ntiles_pad + complement_blocks(A.ui, b, threads));
    switch (g_sweep_waves) {
    case 4:
        if (ref) hipLaunchKernelGGL((k_rbgs_march<T, MGP_MODE_REFERENCE, 4>), g, dim3(256), 0, st, A);
        else hipLaunchKernelGGL((k_rbgs_march<T, MGP_MODE_CONSISTENT, 4>), g, dim3(256), 0, st, A);
        break;
    case 16:
        if (ref) hipLaunchKernelGGL((k_rbgs_march<T, MGP_MODE_REFERENCE, 1>), g, dim3(1024), 0, st, A);
        else hipLaunchKernelGGL((k_rbgs_march<T, MGP_MODE_CONSISTENT, 1>), g, dim3(1024), 0, st, A);
        break;
    default:
        if (ref) hipLaunchKernelGGL((k_rbgs_march<T, MGP_MODE_REFERENCE, 2>), g, dim3(512), 0, st, A);
        else hipLaunchKernelGGL((k_rbgs_march<T, MGP_MODE_CONSISTENT, 2>), g, dim3(512), 0, st, A);
        break;
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? MGP_OK : (int)e;
}


// Residual through the wave kernel; MGP_ERR_ARG when the layout does not
// allow it.  out: where r goes on the box (r itself, or an array of type O
// with the same element layout); the norm reads r outside the box.
template <class T, class O>
int march_residual_to(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r, const mgp_array* out,
                      const int64_t b[6], double h, double* norm, void* stream, bool store) {
    if (!store && !norm) return MGP_ERR_ARG;
    for (int d = 0; d < 3; d++)
        if (u->stride[d] != f->stride[d] || u->stride[d] != r->stride[d] || u->stride[d] != out->stride[d] ||
            u->shape[d] != out->shape[d])
            return MGP_ERR_ARG;
    for (int q = 0; q < 6; q++)
        if (u->halo[q] != f->halo[q] || u->halo[q] != r->halo[q] || u->halo[q] != out->halo[q]) return MGP_ERR_ARG;
    ResidArgs<T, O> A;
    A.u = view_of<T>(*u);
    A.f = view_of<T>(*f);
    A.r = view_of<T>(*r);
    A.out = view_of<O>(*out);
    A.bx0 = b[0], A.bx1 = b[1], A.by0 = b[2], A.by1 = b[3], A.bz0 = b[4], A.bz1 = b[5];
    A.xlo = -u->halo[MGP_WEST], A.xhi = u->shape[0] - 1 + u->halo[MGP_EAST];
    A.ylo = -u->halo[MGP_SOUTH], A.yhi = u->shape[1] - 1 + u->halo[MGP_NORTH];
    A.zlo = -u->halo[MGP_BOTTOM], A.zhi = u->shape[2] - 1 + u->halo[MGP_TOP];
    const int64_t plane_bytes = ((A.yhi - A.ylo) * u->stride[1] + (A.zhi - A.zlo) + 1) * (int64_t)sizeof(T);
    if (plane_bytes >= (int64_t)BUF_OOB) return MGP_ERR_ARG;
    A.hsq = (T)(h * h);  // include/residual.h:23
    A.inv_h2 = (T)(1.0 / (h * h));
    A.norm = (unsigned long long*)norm;
    const int rows = g_resid_rows == 8 || g_resid_rows == 16 ? (int)g_resid_rows : 16;
    const int64_t nx = b[1] - b[0], ny = b[3] - b[2], nz = b[5] - b[4];
    const hipStream_t st = (hipStream_t)stream;
    if (nx > 0 && ny > 0 && nz > 0) {
        A.ytiles = (int)((ny + rows - 1) / rows);
        A.ztiles = (int)((nz + RZ - 1) / RZ);
        const int64_t tiles = (int64_t)A.ytiles * A.ztiles;
        // the norm-only form (nothing stored) streams two arrays and prefers
        // fewer, longer marches; a stored residual keeps the sweep target
        const int64_t tw = store ? g_sweep_target_wgs : target_wgs(g_resid_target);
        int64_t chunks = std::max<int64_t>(1, tw / std::max<int64_t>(tiles, 1));
        chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, nx / std::max<int64_t>(min_chunk_for(box_points(b)), 1)));
        A.xchunk = (nx + chunks - 1) / chunks;
        A.ntiles = (int)(tiles * ((nx + A.xchunk - 1) / A.xchunk));
        A.ntiles_pad = (A.ntiles + 7) / 8 * 8;
        const dim3 g(A.ntiles_pad), blk(64);
        const bool ref = mode == MGP_MODE_REFERENCE;
#define MGP_RES_LAUNCH(RR, NN)                                                                             \
    if (ref) hipLaunchKernelGGL((k_residual_wave<T, O, MGP_MODE_REFERENCE, RR, NN>), g, blk, 0, st, A); \
    else hipLaunchKernelGGL((k_residual_wave<T, O, MGP_MODE_CONSISTENT, RR, NN>), g, blk, 0, st, A);
        if (!store) {
            if (ref) hipLaunchKernelGGL((k_residual_wave<T, O, MGP_MODE_REFERENCE, 16, true, false>), g, blk, 0, st, A);
            else hipLaunchKernelGGL((k_residual_wave<T, O, MGP_MODE_CONSISTENT, 16, true, false>), g, blk, 0, st, A);
        } else if (norm) {
            if (rows == 8) { MGP_RES_LAUNCH(8, true) } else { MGP_RES_LAUNCH(16, true) }
        } else {
            if (rows == 8) { MGP_RES_LAUNCH(8, false) } else { MGP_RES_LAUNCH(16, false) }
        }
#undef MGP_RES_LAUNCH
    }
    if (norm) {
        // the shape points outside the box, whose r is read as stored
        const int64_t n0 = u->shape[0], n1 = u->shape[1], n2 = u->shape[2];
        struct Reg { int64_t x0, x1, y0, y1, z0, z1; };
        const Reg regs[6] = {{0, b[0], 0, n1, 0, n2},       {b[1], n0, 0, n1, 0, n2},
                             {b[0], b[1], 0, b[2], 0, n2},  {b[0], b[1], b[3], n1, 0, n2},
                             {b[0], b[1], b[2], b[3], 0, b[4]}, {b[0], b[1], b[2], b[3], b[5], n2}};
        for (const Reg& g : regs) {
            const int64_t cnt = (g.x1 - g.x0) * (g.y1 - g.y0) * (g.z1 - g.z0);
            if (g.x1 <= g.x0 || g.y1 <= g.y0 || g.z1 <= g.z0) continue;
            const int blocks = (int)std::min<int64_t>(1024, (cnt + 255) / 256);
            hipLaunchKernelGGL(k_norm_region<T>, dim3(blocks), dim3(256), 0, st, A.r, g.x0, g.x1, g.y0, g.y1, g.z0,
                               g.z1, (unsigned long long*)norm);
        }
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? MGP_OK : (int)e;
}

int march_update_residual(int mode, const mgp_array* u, const mgp_array* un, const mgp_array* e, const mgp_array* f,
                          const mgp_array* r, const mgp_array* out, const mgp_faces* faces, const int64_t b[6],
                          double* norm, void* stream) {
    for (int d = 0; d < 3; d++)
        if (u->stride[d] != un->stride[d] || u->stride[d] != f->stride[d] || u->stride[d] != r->stride[d] ||
            u->stride[d] != e->stride[d] || u->stride[d] != out->stride[d] || u->shape[d] != e->shape[d] ||
            u->shape[d] != out->shape[d] || u->shape[d] != un->shape[d])
            return MGP_ERR_ARG;
    for (int q = 0; q < 6; q++)
        if (u->halo[q] != un->halo[q] || u->halo[q] != f->halo[q] || u->halo[q] != r->halo[q] ||
            u->halo[q] != e->halo[q] || u->halo[q] != out->halo[q])
            return MGP_ERR_ARG;
    UpdArgs A;
    A.u = view_of<double>(*u);
    A.un = view_of<double>(*un);
    A.f = view_of<double>(*f);
    A.r = view_of<double>(*r);
    A.e = view_of<float>(*e);
    A.out = view_of<float>(*out);
    for (int q = 0; q < 6; q++)
        A.g[q] = faces->type[q] == MGP_NEUMANN ? static_cast<const double*>(faces->g[q]) : nullptr;
    A.two_h = 2.0 * faces->h;  // include/neumann.h:108
    A.bx0 = b[0], A.bx1 = b[1], A.by0 = b[2], A.by1 = b[3], A.bz0 = b[4], A.bz1 = b[5];
    A.xlo = -u->halo[MGP_WEST], A.xhi = u->shape[0] - 1 + u->halo[MGP_EAST];
    A.ylo = -u->halo[MGP_SOUTH], A.yhi = u->shape[1] - 1 + u->halo[MGP_NORTH];
    A.zlo = -u->halo[MGP_BOTTOM], A.zhi = u->shape[2] - 1 + u->halo[MGP_TOP];
    const int64_t plane_bytes = ((A.yhi - A.ylo) * u->stride[1] + (A.zhi - A.zlo) + 1) * (int64_t)sizeof(double);
    if (plane_bytes >= (int64_t)BUF_OOB) return MGP_ERR_ARG;
    A.hsq = faces->h * faces->h;  // include/residual.h:23
    A.inv_h2 = 1.0 / (faces->h * faces->h);
    A.norm = (unsigned long long*)norm;
    constexpr int R = 16;
    const int64_t nx = b[1] - b[0], ny = b[3] - b[2], nz = b[5] - b[4];
    const hipStream_t st = (hipStream_t)stream;
    A.ytiles = A.ztiles = A.ntiles = A.ntiles_pad = 0;
    A.xchunk = 1;
    if (nx > 0 && ny > 0 && nz > 0) {
        A.ytiles = (int)((ny + R - 1) / R);
        A.ztiles = (int)((nz + RZ - 1) / RZ);
        const int64_t tiles = (int64_t)A.ytiles * A.ztiles;
        int64_t chunks = std::max<int64_t>(1, target_wgs(g_upd_target) / std::max<int64_t>(tiles, 1));
        chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, nx / std::max<int64_t>(min_chunk_for(box_points(b)), 1)));
        A.xchunk = (nx + chunks - 1) / chunks;
        A.ntiles = (int)(tiles * ((nx + A.xchunk - 1) / A.xchunk));
        A.ntiles_pad = (A.ntiles + 7) / 8 * 8;
    }
    const int copy = complement_blocks(A.u, b, 64);
    if (A.ntiles_pad + copy > 0) {
        const dim3 g(A.ntiles_pad + copy), blk(64);
        if (mode == MGP_MODE_REFERENCE)
            hipLaunchKernelGGL((k_update_residual_wave<MGP_MODE_REFERENCE, R>), g, blk, 0, st, A);
        else
            hipLaunchKernelGGL((k_update_residual_wave<MGP_MODE_CONSISTENT, R>), g, blk, 0, st, A);
    }
    if (norm) {
        // the shape points outside the box, whose r is read as stored
        const int64_t n0 = u->shape[0], n1 = u->shape[1], n2 = u->shape[2];
        struct Reg { int64_t x0, x1, y0, y1, z0, z1; };
        const Reg regs[6] = {{0, b[0], 0, n1, 0, n2},       {b[1], n0, 0, n1, 0, n2},
                             {b[0], b[1], 0, b[2], 0, n2},  {b[0], b[1], b[3], n1, 0, n2},
                             {b[0], b[1], b[2], b[3], 0, b[4]}, {b[0], b[1], b[2], b[3], b[5], n2}};
        for (const Reg& g : regs) {
            const int64_t cnt = (g.x1 - g.x0) * (g.y1 - g.y0) * (g.z1 - g.z0);
            if (g.x1 <= g.x0 || g.y1 <= g.y0 || g.z1 <= g.z0) continue;
            const int blocks = (int)std::min<int64_t>(1024, (cnt + 255) / 256);
            hipLaunchKernelGGL(k_norm_region<double>, dim3(blocks), dim3(256), 0, st, A.r, g.x0, g.x1, g.y0, g.y1,
                               g.z0, g.z1, (unsigned long long*)norm);
        }
    }
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MGP_OK : (int)err;
}

// rows per wave tile of the hi + lo update + residual: 8 or 16
int64_t g_hilo_rows = 16;

int march_hilo_residual(const mgp_array* u, const mgp_array* un, const mgp_array* lo, const mgp_array* ln,
                        const mgp_array* e, const mgp_array* f, const mgp_array* r, const mgp_array* out,
                        const mgp_faces* faces, const int64_t b[6], double* norm, void* stream) {
    const bool upd = e != nullptr;
    const mgp_array* fl[] = {lo, upd ? ln : lo, upd ? e : lo};
    const mgp_array* dl[] = {un ? un : u, f, r};
    for (int d = 0; d < 3; d++) {
        for (const mgp_array* a : fl)
            if (a->stride[d] != u->stride[d] || a->shape[d] != u->shape[d]) return MGP_ERR_ARG;
        for (const mgp_array* a : dl)
            if (a->stride[d] != u->stride[d] || a->shape[d] != u->shape[d]) return MGP_ERR_ARG;
        if (out->stride[d] != u->stride[d] || out->shape[d] != u->shape[d]) return MGP_ERR_ARG;
    }
    for (int q = 0; q < 6; q++) {
        for (const mgp_array* a : fl)
            if (a->halo[q] != u->halo[q]) return MGP_ERR_ARG;
        for (const mgp_array* a : dl)
            if (a->halo[q] != u->halo[q]) return MGP_ERR_ARG;
        if (out->halo[q] != u->halo[q]) return MGP_ERR_ARG;
    }
    HiloArgs A;
    A.u = view_of<double>(*u);
    A.un = view_of<double>(upd ? *un : *u);
    A.f = view_of<double>(*f);
    A.r = view_of<double>(*r);
    A.lo = view_of<float>(*lo);
    A.ln = view_of<float>(upd ? *ln : *lo);
    A.e = view_of<float>(upd ? *e : *lo);
    A.out64 = upd ? nullptr : view_of<double>(*out).p;
    A.out32 = upd ? view_of<float>(*out).p : nullptr;
    for (int q = 0; q < 6; q++)
        A.g[q] = faces->type[q] == MGP_NEUMANN ? static_cast<const double*>(faces->g[q]) : nullptr;
    A.two_h = 2.0 * faces->h;  // include/neumann.h:108
    A.bx0 = b[0], A.bx1 = b[1], A.by0 = b[2], A.by1 = b[3], A.bz0 = b[4], A.bz1 = b[5];
    A.xlo = -u->halo[MGP_WEST], A.xhi = u->shape[0] - 1 + u->halo[MGP_EAST];
    A.ylo = -u->halo[MGP_SOUTH], A.yhi = u->shape[1] - 1 + u->halo[MGP_NORTH];
    A.zlo = -u->halo[MGP_BOTTOM], A.zhi = u->shape[2] - 1 + u->halo[MGP_TOP];
    const int64_t plane_bytes = ((A.yhi - A.ylo) * u->stride[1] + (A.zhi - A.zlo) + 1) * (int64_t)sizeof(double);
    if (plane_bytes >= (int64_t)BUF_OOB) return MGP_ERR_ARG;
    A.inv_h2 = 1.0 / (faces->h * faces->h);
    A.norm = (unsigned long long*)norm;
    const int R = g_hilo_rows == 8 ? 8 : 16;
    const int64_t nx = b[1] - b[0], ny = b[3] - b[2], nz = b[5] - b[4];
    const hipStream_t st = (hipStream_t)stream;
    A.ytiles = A.ztiles = A.ntiles = A.ntiles_pad = 0;
    A.xchunk = 1;
    if (nx > 0 && ny > 0 && nz > 0) {
        A.ytiles = (int)((ny + R - 1) / R);
        A.ztiles = (int)((nz + RZ - 1) / RZ);
        const int64_t tiles = (int64_t)A.ytiles * A.ztiles;
        int64_t chunks = std::max<int64_t>(1, target_wgs(g_upd_target) / std::max<int64_t>(tiles, 1));
        chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, nx / std::max<int64_t>(min_chunk_for(box_points(b)), 1)));
        A.xchunk = (nx + chunks - 1) / chunks;
        A.ntiles = (int)(tiles * ((nx + A.xchunk - 1) / A.xchunk));
        A.ntiles_pad = (A.ntiles + 7) / 8 * 8;
    }
    const int copy = upd ? complement_blocks(A.u, b, 64) : 0;
    if (A.ntiles_pad + copy > 0) {
        const dim3 g(A.ntiles_pad + copy), blk(64);
        if (upd) {
            if (R == 8) hipLaunchKernelGGL((k_hilo_residual_wave<8, true>), g, blk, 0, st, A);
            else hipLaunchKernelGGL((k_hilo_residual_wave<16, true>), g, blk, 0, st, A);
        } else {
            if (R == 8) hipLaunchKernelGGL((k_hilo_residual_wave<8, false>), g, blk, 0, st, A);
            else hipLaunchKernelGGL((k_hilo_residual_wave<16, false>), g, blk, 0, st, A);
        }
    }
    if (norm) {
        // the shape points outside the box, whose r is read as stored (FP64 r)
        const int64_t n0 = u->shape[0], n1 = u->shape[1], n2 = u->shape[2];
        struct Reg { int64_t x0, x1, y0, y1, z0, z1; };
        const Reg regs[6] = {{0, b[0], 0, n1, 0, n2},       {b[1], n0, 0, n1, 0, n2},
                             {b[0], b[1], 0, b[2], 0, n2},  {b[0], b[1], b[3], n1, 0, n2},
                             {b[0], b[1], b[2], b[3], 0, b[4]}, {b[0], b[1], b[2], b[3], b[5], n2}};
        for (const Reg& g : regs) {
            const int64_t cnt = (g.x1 - g.x0) * (g.y1 - g.y0) * (g.z1 - g.z0);
            if (g.x1 <= g.x0 || g.y1 <= g.y0 || g.z1 <= g.z0) continue;
            const int blocks = (int)std::min<int64_t>(1024, (cnt + 255) / 256);
            hipLaunchKernelGGL(k_norm_region<double>, dim3(blocks), dim3(256), 0, st, A.r, g.x0, g.x1, g.y0, g.y1,
                               g.z0, g.z1, (unsigned long long*)norm);
        }
    }
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MGP_OK : (int)err;
}

template <class T>
int march_residual(int mode, const mgp_array* u, const mgp_array* f, const mgp_array* r, const int64_t b[6],
                   double h, double* norm, void* stream) {
    return march_residual_to<T, T>(mode, u, f, r, r, b, h, norm, stream, true);
}

int64_t g_resid_kernel = 2;  // 1: per-point kernels, 2: wave tiles
int64_t g_resid_rows = 16;   // register rows per residual wave tile: 8 or 16

template int march_residual<double>(int, const mgp_array*, const mgp_array*, const mgp_array*, const int64_t[6],
                                    double, double*, void*);
template int march_residual<float>(int, const mgp_array*, const mgp_array*, const mgp_array*, const int64_t[6],
                                   double, double*, void*);
template int march_residual_to<double, float>(int, const mgp_array*, const mgp_array*, const mgp_array*,
                                              const mgp_array*, const int64_t[6], double, double*, void*, bool);
template int march_residual_to<double, double>(int, const mgp_array*, const mgp_array*, const mgp_array*,
                                               const mgp_array*, const int64_t[6], double, double*, void*, bool);
template int march_residual_to<float, float>(int, const mgp_array*, const mgp_array*, const mgp_array*,
                                             const mgp_array*, const int64_t[6], double, double*, void*, bool);


// u += P(coarse) through wave tiles; MGP_ERR_ARG when the planes exceed the
// 32-bit buffer range
template <class T>
int march_prolong_add(const mgp_array* coarse, const mgp_array* fine, void* stream) {
    ProlongArgs<T> A;
    A.c = view_of<T>(*coarse);
    A.u = view_of<T>(*fine);
    const int64_t fp = ((A.u.n1 - 1) * A.u.s1 + A.u.n2) * (int64_t)sizeof(T);
    const int64_t cp = ((A.c.n1 - 1) * A.c.s1 + A.c.n2) * (int64_t)sizeof(T);
    if (fp >= (int64_t)BUF_OOB || cp >= (int64_t)BUF_OOB) return MGP_ERR_ARG;
    constexpr int R = 16;
    const int64_t nx = A.u.n0;
    A.ytiles = (int)((A.u.n1 + R - 1) / R);
    A.ztiles = (int)((A.u.n2 + 63) / 64);
    const int64_t tiles = (int64_t)A.ytiles * A.ztiles;
    int64_t chunks = std::max<int64_t>(1, target_wgs(g_prolong_target) / std::max<int64_t>(tiles, 1));
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, nx / std::max<int64_t>(min_chunk_for(A.u.n0 * A.u.n1 * A.u.n2), 2)));
    int64_t xchunk = (nx + chunks - 1) / chunks;
    xchunk += xchunk & 1;  // chunks start on even planes
    A.xchunk = xchunk;
    A.ntiles = (int)(tiles * ((nx + xchunk - 1) / xchunk));
    A.ntiles_pad = (A.ntiles + 7) / 8 * 8;
    hipLaunchKernelGGL((k_prolong_add_wave<T, R>), dim3(A.ntiles_pad), dim3(64), 0, (hipStream_t)stream, A);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? MGP_OK : (int)e;
}
template int march_prolong_add<double>(const mgp_array*, const mgp_array*, void*);
template int march_prolong_add<float>(const mgp_array*, const mgp_array*, void*);
int64_t g_prolong_kernel = 2;  // 1: per-point kernel, 2: wave tiles


// coarse f = FW(f - L u) in one pass (consistent mode); MGP_ERR_ARG when the
// layouts do not allow it (then the caller runs the residual and the
// restriction separately)
template <class T>
int march_residual_fw(const mgp_array* u, const mgp_array* f, const mgp_array* r, const mgp_array* cf,
                      const int64_t b[6], double h, void* stream) {
    for (int d = 0; d < 3; d++)
        if (u->stride[d] != f->stride[d] || u->stride[d] != r->stride[d] || (u->shape[d] - 1) / 2 + 1 != cf->shape[d])
            return MGP_ERR_ARG;
    for (int q = 0; q < 6; q++)
        if (u->halo[q] != f->halo[q] || u->halo[q] != r->halo[q]) return MGP_ERR_ARG;
    RFArgs<T> A;
    A.u = view_of<T>(*u);
    A.f = view_of<T>(*f);
    A.r = view_of<T>(*r);
    A.c = view_of<T>(*cf);
    A.bx0 = b[0], A.bx1 = b[1], A.by0 = b[2], A.by1 = b[3], A.bz0 = b[4], A.bz1 = b[5];
    A.xlo = -u->halo[MGP_WEST], A.xhi = u->shape[0] - 1 + u->halo[MGP_EAST];
    A.ylo = -u->halo[MGP_SOUTH], A.yhi = u->shape[1] - 1 + u->halo[MGP_NORTH];
    A.zlo = -u->halo[MGP_BOTTOM], A.zhi = u->shape[2] - 1 + u->halo[MGP_TOP];
    const int64_t plane_bytes = ((A.yhi - A.ylo) * u->stride[1] + (A.zhi - A.zlo) + 1) * (int64_t)sizeof(T);
    const int64_t cplane = ((cf->shape[1] - 1) * cf->stride[1] + cf->shape[2]) * (int64_t)sizeof(T);
    if (plane_bytes >= (int64_t)BUF_OOB || cplane >= (int64_t)BUF_OOB) return MGP_ERR_ARG;
    A.hsq = (T)(h * h);  // include/residual.h:23
    A.inv_h2 = (T)(1.0 / (h * h));
    // measured at 1025 x 1025 x 513: 6 rows for both types, two planes in
    // flight for FP32 (FP64 at 6 rows and depth 2 spills)
    const int RJ = g_fw_rows == 4 ? 4 : 6;
    const int depth = g_fw_depth == 1 || g_fw_depth == 2 ? (int)g_fw_depth : (sizeof(T) == 8 ? 1 : 2);
    const int64_t c0 = cf->shape[0];
    A.ytiles = (int)((cf->shape[1] + RJ - 1) / RJ);
    A.ztiles = (int)((cf->shape[2] + FWK - 1) / FWK);
    const int64_t tiles = (int64_t)A.ytiles * A.ztiles;
    int64_t chunks = std::max<int64_t>(1, target_wgs(g_fw_target) / std::max<int64_t>(tiles, 1));
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, c0 / 2));
    A.ichunk = (c0 + chunks - 1) / chunks;
    A.ntiles = (int)(tiles * ((c0 + A.ichunk - 1) / A.ichunk));
    A.ntiles_pad = (A.ntiles + 7) / 8 * 8;
    const hipStream_t st = (hipStream_t)stream;
    if (RJ == 6) {
        if (depth == 2) hipLaunchKernelGGL((k_resid_fw_wave<T, 6, 2>), dim3(A.ntiles_pad), dim3(64), 0, st, A);
        else hipLaunchKernelGGL((k_resid_fw_wave<T, 6, 1>), dim3(A.ntiles_pad), dim3(64), 0, st, A);
    } else {
        if (depth == 2) hipLaunchKernelGGL((k_resid_fw_wave<T, 4, 2>), dim3(A.ntiles_pad), dim3(64), 0, st, A);
        else hipLaunchKernelGGL((k_resid_fw_wave<T, 4, 1>), dim3(A.ntiles_pad), dim3(64), 0, st, A);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? MGP_OK : (int)e;
}
template int march_residual_fw<double>(const mgp_array*, const mgp_array*, const mgp_array*, const mgp_array*,
                                       const int64_t[6], double, void*);
template int march_residual_fw<float>(const mgp_array*, const mgp_array*, const mgp_array*, const mgp_array*,
                                      const int64_t[6], double, void*);

int64_t g_fw_rows = 0;   // coarse rows per residual + full-weighting wave tile: 4 or 6; 0: 6
int64_t g_fw_depth = 0;  // fine planes of loads in flight per wave: 1 or 2; 0: 1 (FP64), 2 (FP32)

// Launch-shape knobs of the one-pass sweep (set through mgp_tuning_set).
int64_t g_sweep2_cpt = 15;          // two-sweep pass: least x chunks per (y, z) tile
int64_t g_sweep_target_wgs = 16384;  // workgroups (waves, kernel 2) to aim for per launch
int64_t g_sweep_min_chunk = 4;      // shortest x march per workgroup / wave: coarse levels are
                                    // latency-bound, so they get many short marches
int64_t g_sweep_waves = 8;          // waves per workgroup: 4, 8 or 16
int64_t g_sweep_kernel = 2;         // 1: workgroup tiles with LDS, 2: wave-private tiles
int64_t g_sweep_rows = 0;           // register rows per wave tile: 8, 12, 16; 0: by type
int64_t g_sweep_depth = 1;          // planes of loads in flight per wave: 1 or 2

// x-chunk length: enough workgroups to keep every CU busy with several
// marches, but chunks long enough that the two warm-up planes per chunk stay
// cheap
int64_t march_xchunk(const int64_t b[6]) {
    const int64_t tiles = ((b[3] - b[2] + TY - 1) / TY) * ((b[5] - b[4] + TZ - 1) / TZ);
    const int64_t nx = b[1] - b[0];
    int64_t chunks = std::max<int64_t>(1, g_sweep_target_wgs / std::max<int64_t>(tiles, 1));
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, nx / std::max<int64_t>(min_chunk_for(box_points(b)), 1)));
    return (nx + chunks - 1) / chunks;
}

template int march_sweep<double>(int, const mgp_array*, const mgp_array*, const mgp_array*, const mgp_faces*,
                                 const int64_t[6], double, int64_t, void*);
template int march_sweep<float>(int, const mgp_array*, const mgp_array*, const mgp_array*, const mgp_faces*,
                                const int64_t[6], double, int64_t, void*);

int64_t g_sweep_pair = 1;   // relax_level runs sweeps in pairs through the two-sweep pass
int64_t g_sweep2_rows = 0;  // register rows per two-sweep tile: FP64 12 or 16, FP32 16, 20 or 24; 0: 16
                            // (measured best for both at 1025 x 1025 x 513)
// consistent-mode two-sweep stages evaluate row pairs packed by colour (bit 0
// FP64, bit 1 FP32).  1025 x 1025 x 513: FP64 4.53 ms against 4.92 ms per pass,
// with the correction inside 5.71 against 5.80 ms; FP32 2.29 against 2.48 ms
int64_t g_sweep2_pack = 3;
int64_t g_sweep2_waves = 6800;  // least waves per two-sweep launch (x-chunks added to reach it)
int64_t g_sweep2_prolong = 2;  // the cycles add the prolongation inside the first post-smoothing pass
                               // (1: FP64 levels, 2: FP32 levels too)

template <class T>
int march_sweep2(int mode, const mgp_array* ui, const mgp_array* uo, const mgp_array* f, const mgp_faces* faces,
                 const int64_t b[6], double omega, void* stream, const mgp_array* coarse, bool zero_in) {
    MarchArgs<T> A;
    A.ui = view_of<T>(*ui);
    A.uo = view_of<T>(*uo);
    A.f = view_of<T>(*f);
    A.c = coarse ? view_of<T>(*coarse) : A.ui;
    // planes addressed through buffer descriptors with a 32-bit range
    const int64_t plane_bytes = ((A.ui.n1 - 1) * A.ui.s1 + A.ui.n2) * (int64_t)sizeof(T);
    if (A.ui.n0 < 3 || plane_bytes >= (int64_t)BUF_OOB) return MGP_ERR_ARG;
    if (coarse) {
        const int64_t cplane = ((A.c.n1 - 1) * A.c.s1 + A.c.n2) * (int64_t)sizeof(T);
        if (mode != MGP_MODE_CONSISTENT || cplane >= (int64_t)BUF_OOB || zero_in) return MGP_ERR_ARG;
    }
    if (zero_in && mode != MGP_MODE_CONSISTENT) return MGP_ERR_ARG;
    for (int q = 0; q < 6; q++) A.g[q] = faces->type[q] == MGP_NEUMANN ? static_cast<const T*>(faces->g[q]) : nullptr;
    A.two_h = (T)(2.0 * faces->h);  // include/neumann.h:108
    A.bx0 = b[0], A.bx1 = b[1], A.by0 = b[2], A.by1 = b[3], A.bz0 = b[4], A.bz1 = b[5];
    A.lin0 = ui->halo[MGP_WEST] * ui->stride[0] + ui->halo[MGP_SOUTH] * ui->stride[1] + ui->halo[MGP_BOTTOM];
    const T om = (T)omega;  // include/gaussseidel.h:50-52
    A.one_omega = (T)(1.0 - om);
    A.omega_sixth = (T)(om / 6.0);
    A.hsq = (T)(faces->h * faces->h);
    int rows = (int)g_sweep2_rows;
    if (sizeof(T) == 8 && rows != 12 && rows != 16) rows = 16;
    if (sizeof(T) == 4 && rows != 16 && rows != 20 && rows != 24) rows = 16;
    if (coarse || zero_in) rows = 16;  // the one tile height with the prolongation or zero input built in
    A.ytiles = (int)((b[3] - b[2] + rows - 9) / (rows - 8));
    A.ztiles = (int)((b[5] - b[4] + WZ2 - 1) / WZ2);
    const int64_t tiles = (int64_t)A.ytiles * A.ztiles;
    const int64_t nx = b[1] - b[0];
    // x chunks per (y, z) tile (each chunk spends 6 warm-up steps): at least
    // sweep2.chunks_per_tile, enough for sweep2.target_waves waves on coarser levels,
    // and an odd count.  Measured, FP64: 1025 x 1025 x 513 with 13, 15, 17
    // chunks 4.94-5.01 ms, with 12, 14, 16, 18 chunks 5.19-5.21 ms;
    // 513 x 513 x 257 with 21 chunks 0.82 ms, with 16 0.89 ms; 257 x 257 x 129
    // and below flat.
    const int64_t tl = std::max<int64_t>(tiles, 1);
    int64_t chunks = std::max<int64_t>(g_sweep2_cpt, (g_sweep2_waves + tl - 1) / tl) | 1;
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, nx / std::max<int64_t>(min_chunk_for(box_points(b)), 1)));
    A.xchunk = (nx + chunks - 1) / chunks;
    A.ntiles = (int)(tiles * ((nx + A.xchunk - 1) / A.xchunk));
    A.ntiles_pad = (A.ntiles + 7) / 8 * 8;
    const dim3 g(A.ntiles_pad + complement_blocks(A.ui, b, 64));
    const hipStream_t st = (hipStream_t)stream;
    const bool pack = (g_sweep2_pack & (sizeof(T) == 8 ? 1 : 2)) && mode == MGP_MODE_CONSISTENT && rows == 16;
    if (coarse) {
        const int py = (int)(b[2] & 1), pz = (int)(b[4] & 1);  // parities of y0 - 4 and z0 - 4
        if (pack) {
            if (py == 0 && pz == 0) hipLaunchKernelGGL((k_rbgs2p_wave<T, 16, 0, 0, true>), g, dim3(64), 0, st, A);
            else if (py == 0) hipLaunchKernelGGL((k_rbgs2p_wave<T, 16, 0, 1, true>), g, dim3(64), 0, st, A);
            else if (pz == 0) hipLaunchKernelGGL((k_rbgs2p_wave<T, 16, 1, 0, true>), g, dim3(64), 0, st, A);
            else hipLaunchKernelGGL((k_rbgs2p_wave<T, 16, 1, 1, true>), g, dim3(64), 0, st, A);
        } else {
            if (py == 0 && pz == 0) hipLaunchKernelGGL((k_rbgs2p_wave<T, 16, 0, 0>), g, dim3(64), 0, st, A);
            else if (py == 0) hipLaunchKernelGGL((k_rbgs2p_wave<T, 16, 0, 1>), g, dim3(64), 0, st, A);
            else if (pz == 0) hipLaunchKernelGGL((k_rbgs2p_wave<T, 16, 1, 0>), g, dim3(64), 0, st, A);
            else hipLaunchKernelGGL((k_rbgs2p_wave<T, 16, 1, 1>), g, dim3(64), 0, st, A);
        }
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? MGP_OK : (int)e;
    }
    if (zero_in) {
        if (pack) hipLaunchKernelGGL((k_rbgs2z_wave<T, 16, true>), g, dim3(64), 0, st, A);
        else hipLaunchKernelGGL((k_rbgs2z_wave<T, 16, false>), g, dim3(64), 0, st, A);
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? MGP_OK : (int)e;
    }
    // other row counts, packed, f in LDS (sweep2.rows): FP64 12 rows and FP32 20
    // rows run two waves per SIMD, and take what 16 rows at one wave take at
    // 1025 x 1025 x 513 (4.54 against 4.53 ms, 2.32 against 2.30 ms); FP32 24
    // rows 2.65 ms
    if (rows != 16 && (g_sweep2_pack & (sizeof(T) == 8 ? 1 : 2)) && mode == MGP_MODE_CONSISTENT) {
        if constexpr (sizeof(T) == 8) {
            hipLaunchKernelGGL((k_rbgs2f_wave<T, 12>), g, dim3(64), 0, st, A);
        } else {
            if (rows == 24) hipLaunchKernelGGL((k_rbgs2_wave<T, MGP_MODE_CONSISTENT, 24, true, true>), g, dim3(64), 0, st, A);
            else hipLaunchKernelGGL((k_rbgs2f_wave<T, 20>), g, dim3(64), 0, st, A);
        }
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? MGP_OK : (int)e;
    }
    if (pack) {
        hipLaunchKernelGGL((k_rbgs2_wave<T, MGP_MODE_CONSISTENT, 16, true>), g, dim3(64), 0, st, A);
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? MGP_OK : (int)e;
    }
    const bool ref = mode == MGP_MODE_REFERENCE;
#define MGP_WAVE2_LAUNCH(RR)                                                                            \
    if (ref) hipLaunchKernelGGL((k_rbgs2_wave<T, MGP_MODE_REFERENCE, RR>), g, dim3(64), 0, st, A); \
    else hipLaunchKernelGGL((k_rbgs2_wave<T, MGP_MODE_CONSISTENT, RR>), g, dim3(64), 0, st, A);
    if constexpr (sizeof(T) == 8) {
        if (rows == 12) { MGP_WAVE2_LAUNCH(12) } else { MGP_WAVE2_LAUNCH(16) }
    } else {
        if (rows == 16) { MGP_WAVE2_LAUNCH(16) } else if (rows == 24) { MGP_WAVE2_LAUNCH(24) } else { MGP_WAVE2_LAUNCH(20) }
    }
#undef MGP_WAVE2_LAUNCH
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? MGP_OK : (int)e;
}
template int march_sweep2<double>(int, const mgp_array*, const mgp_array*, const mgp_array*, const mgp_faces*,
                                  const int64_t[6], double, void*, const mgp_array*, bool);
template int march_sweep2<float>(int, const mgp_array*, const mgp_array*, const mgp_array*, const mgp_faces*,
                                 const int64_t[6], double, void*, const mgp_array*, bool);

}  // namespace mgp
