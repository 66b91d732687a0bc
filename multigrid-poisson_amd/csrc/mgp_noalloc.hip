// V- and F-cycles without device allocation (mgp_cycle_noalloc_f64).  No
// counterpart in the reference; for callers that cannot call hipMalloc (code
// under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The caller passes the finest level's u and f and its boundary faces; the
// finest level's residual and second u buffer and every coarser level (u, its
// second buffer, f, r and zero Neumann face data) are carved out of one
// file-scope device array, `scratch`, of MGP_NOALLOC_SCRATCH_BYTES per GPU.
// Host code never uses that array's name as an address: its device address is
// taken once per device with hipGetSymbolAddress and kept per device.  The
// cycle runs the same kernels in the same order as PoissonSolver::cycle of a
// consistent-mode FP64 Gauss-Seidel / full-weighting solver (libpoisson.h,
// vcycle.h, fcycle.h, gaussseidel.h of this library), so u ends with the same
// bits.  Every call clears what it reads before writing it (the residual
// arrays, the coarse face data, each coarse u when the cycle descends); calls
// that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "mgp.h"

namespace {

__device__ __align__(256) unsigned char scratch[MGP_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_base[MAX_DEVICES] = {};

// device address of this GPU's scratch array (taken on the first call on it)
int scratch_base(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= MAX_DEVICES) return MGP_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_base[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return MGP_ERR_STATE;
        g_base[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_base[dev];
    return MGP_OK;
}

struct Level {
    mgp_array u, alt, f, r;
    mgp_faces F;
    int64_t box[6];
    int64_t points;
};

// the solver's array layout for a level: one halo layer on each Neumann face
mgp_array layout(const int64_t n[3], const int32_t type[6]) {
    mgp_array a;
    std::memset(&a, 0, sizeof(a));
    for (int d = 0; d < 3; d++) a.shape[d] = n[d];
    for (int q = 0; q < 6; q++) a.halo[q] = type[q] == MGP_NEUMANN ? 1 : 0;
    a.stride[2] = 1;
    a.stride[1] = n[2] + a.halo[MGP_TOP] + a.halo[MGP_BOTTOM];
    a.stride[0] = (n[1] + a.halo[MGP_NORTH] + a.halo[MGP_SOUTH]) * a.stride[1];
    return a;
}
int64_t elements(const mgp_array& a) { return (a.shape[0] + a.halo[MGP_EAST] + a.halo[MGP_WEST]) * a.stride[0]; }
int64_t face_points(const int64_t n[3], int q) {
    if (q == MGP_EAST || q == MGP_WEST) return n[1] * n[2];
    if (q == MGP_NORTH || q == MGP_SOUTH) return n[0] * n[2];
    return n[0] * n[1];
}

struct Cycle {
    std::vector<Level> L;
    double omega;
    int64_t nsmooth, coarse;
    int64_t pair, pair_min;
    void* st;
    int rc = MGP_OK;

    bool ok(int r) {
        if (rc == MGP_OK && r != MGP_OK) rc = r;
        return rc == MGP_OK;
    }
    static void swap(Level& d) { std::swap(d.u.data, d.alt.data); }

    // GaussSeidel::relax_level (consistent mode, fused sweeps): one LDS launch
    // where the level fits, else pairs of sweeps on large levels, then single
    // sweeps, each out of place into the other buffer
    void relax(int l, int64_t n) {
        Level& d = L[l];
        if (n < 1 || rc != MGP_OK) return;
        mgp_array uo = d.u;
        if (n & 1) uo.data = d.alt.data;
        if (mgp_rbgs_sweeps_f64(MGP_MODE_CONSISTENT, &d.u, &uo, &d.f, &d.F, d.box, omega, n, st) == MGP_OK) {
            if (n & 1) swap(d);
            return;
        }
        int64_t s = 0;
        const bool pairs = pair && d.points >= pair_min;
        for (; pairs && s + 2 <= n; s += 2) {
            uo = d.u;
            uo.data = d.alt.data;
            const int r = mgp_rbgs_sweep2_f64(MGP_MODE_CONSISTENT, &d.u, &uo, &d.f, &d.F, d.box, omega, st);
            if (r == MGP_ERR_ARG) break;
            if (!ok(r)) return;
            swap(d);
        }
        for (; s < n; s++) {
            uo = d.u;
            uo.data = d.alt.data;
            if (!ok(mgp_rbgs_sweep_f64(MGP_MODE_CONSISTENT, &d.u, &uo, &d.f, &d.F, d.box, omega, st))) return;
            swap(d);
        }
    }
    // Neumann ghosts of the current u, in the boundary refresh order E W N S T B
    void refresh(int l) {
        Level& d = L[l];
        for (int q : {MGP_EAST, MGP_WEST, MGP_NORTH, MGP_SOUTH, MGP_TOP, MGP_BOTTOM})
            if (d.F.type[q] == MGP_NEUMANN && !ok(mgp_bc_neumann_f64(&d.u, q, d.F.g[q], d.F.h, st))) return;
    }
    // residual and full weighting into the coarse f, coarse u = 0
    void descend(int l) {
        Level& d = L[l];
        Level& c = L[l + 1];
        refresh(l);
        if (rc != MGP_OK) return;
        if (mgp_residual_restrict_fw_f64(MGP_MODE_CONSISTENT, &d.u, &d.f, &d.r, &c.f, d.box, d.F.h, st) != MGP_OK) {
            refresh(l);
            if (!ok(mgp_residual_f64(MGP_MODE_CONSISTENT, &d.u, &d.f, &d.r, d.box, d.F.h, nullptr, st))) return;
            if (!ok(mgp_restrict_fw_f64(&d.r, &c.f, st))) return;
        }
        ok(mgp_fill_f64(&c.u, 0.0, st));
    }
    void ascend(int l) {
        if (rc == MGP_OK) ok(mgp_prolong_add_f64(&L[l + 1].u, &L[l].u, st));
    }
    void vcycle(int l) {
        const int levels = (int)L.size();
        if (l >= levels - 1) {
            relax(l, coarse > 0 ? coarse : nsmooth);
            return;
        }
        relax(l, nsmooth);
        descend(l);
        vcycle(l + 1);
        ascend(l);
        relax(l, nsmooth);
    }
    void fcycle(int l) {
        const int levels = (int)L.size();
        relax(l, nsmooth);
        descend(l);
        if (l < levels - 2) fcycle(l + 1);
        vcycle(l + 1);
        ascend(l);
        if (l == 0) vcycle(0);
    }
};

bool same_layout(const mgp_array* a, const mgp_array& b) {
    for (int d = 0; d < 3; d++)
        if (a->shape[d] != b.shape[d] || a->stride[d] != b.stride[d]) return false;
    for (int q = 0; q < 6; q++)
        if (a->halo[q] != b.halo[q]) return false;
    return a->data != nullptr;
}

}  // namespace

extern "C" int mgp_cycle_noalloc_f64(const char* cycle, const mgp_array* u, const mgp_array* f,
                                     const mgp_faces* faces, int64_t levels, int64_t ncycles, int64_t nsmooth,
                                     double omega, int64_t coarse_sweeps, double* norm, void* stream) {
    if (!cycle || !u || !f || !faces || levels < 1 || levels > 30 || ncycles < 0 || nsmooth < 0 ||
        coarse_sweeps < 0 || !(faces->h > 0.0) || !std::isfinite(omega))
        return MGP_ERR_ARG;
    const bool v = !std::strcmp(cycle, "Vcycle") || !std::strcmp(cycle, "vcycle");
    const bool fc = !std::strcmp(cycle, "Fcycle") || !std::strcmp(cycle, "fcycle");
    if (!(v || fc) || (fc && levels < 2)) return MGP_ERR_ARG;
    for (int q = 0; q < 6; q++) {
        if (faces->type[q] != MGP_DIRICHLET && faces->type[q] != MGP_NEUMANN) return MGP_ERR_ARG;
        if (faces->type[q] == MGP_NEUMANN && !faces->g[q]) return MGP_ERR_ARG;
    }
    const int64_t n0[3] = {u->shape[0], u->shape[1], u->shape[2]};
    const mgp_array fine = layout(n0, faces->type);
    if (!same_layout(u, fine) || !same_layout(f, fine)) return MGP_ERR_ARG;
    // level shapes ((n >> l) + 1 per axis, n = 2^(levels-1) m + 1) and the
    // scratch they take, 256-byte aligned pieces
    const int64_t step = (int64_t)1 << (levels - 1);
    for (int d = 0; d < 3; d++)
        if (n0[d] < 3 || (n0[d] - 1) % step != 0 || (n0[d] - 1) / step < 1) return MGP_ERR_ARG;
    auto piece = [](int64_t bytes) { return (bytes + 255) / 256 * 256; };
    int64_t need = 0;
    for (int64_t l = 0; l < levels; l++) {
        const int64_t n[3] = {l ? (n0[0] >> l) + 1 : n0[0], l ? (n0[1] >> l) + 1 : n0[1], l ? (n0[2] >> l) + 1 : n0[2]};
        const int64_t bytes = piece(elements(layout(n, faces->type)) * (int64_t)sizeof(double));
        need += (l ? 4 : 2) * bytes;
        if (l)
            for (int q = 0; q < 6; q++)
                if (faces->type[q] == MGP_NEUMANN) need += piece(face_points(n, q) * (int64_t)sizeof(double));
    }
    if (need > (int64_t)MGP_NOALLOC_SCRATCH_BYTES) return MGP_ERR_ARG;
    unsigned char* base = nullptr;
    int rc = scratch_base(&base);
    if (rc != MGP_OK) return rc;
    Cycle c;
    c.omega = omega < 0.0 ? 8.0 / 9.0 : omega;  // include/gaussseidel.h:97-99
    c.nsmooth = nsmooth;
    c.coarse = coarse_sweeps;
    c.st = stream;
    if (mgp_tuning_get("sweep.pair", &c.pair) != MGP_OK) c.pair = 1;
    if (mgp_tuning_get("sweep.pair_min_points", &c.pair_min) != MGP_OK) c.pair_min = 0;
    unsigned char* next = base;
    auto take = [&](int64_t bytes) {
        unsigned char* p = next;
        next += piece(bytes);
        return p;
    };
    for (int64_t l = 0; l < levels; l++) {
        Level d;
        const int64_t n[3] = {l ? (n0[0] >> l) + 1 : n0[0], l ? (n0[1] >> l) + 1 : n0[1], l ? (n0[2] >> l) + 1 : n0[2]};
        const mgp_array a = layout(n, faces->type);
        const int64_t bytes = elements(a) * (int64_t)sizeof(double);
        d.u = d.alt = d.f = d.r = a;
        if (l == 0) {
            d.u.data = u->data;
            d.f.data = f->data;
        } else {
            d.u.data = take(bytes);
            d.f.data = take(bytes);
        }
        d.alt.data = take(bytes);
        d.r.data = take(bytes);
        d.F = *faces;
        d.F.h = std::ldexp(faces->h, (int)l);
        for (int q = 0; q < 6; q++) {
            if (l == 0 || faces->type[q] != MGP_NEUMANN) continue;
            mgp_array g;
            std::memset(&g, 0, sizeof(g));
            g.shape[0] = face_points(n, q);
            g.shape[1] = g.shape[2] = 1;
            g.stride[0] = 1;
            g.stride[1] = g.stride[2] = 1;
            g.data = take(face_points(n, q) * (int64_t)sizeof(double));
            d.F.g[q] = g.data;
            if (!c.ok(mgp_fill_f64(&g, 0.0, stream))) return c.rc;  // homogeneous coarse data
        }
        d.box[0] = faces->type[MGP_WEST] == MGP_DIRICHLET;
        d.box[1] = n[0] - (faces->type[MGP_EAST] == MGP_DIRICHLET);
        d.box[2] = faces->type[MGP_SOUTH] == MGP_DIRICHLET;
        d.box[3] = n[1] - (faces->type[MGP_NORTH] == MGP_DIRICHLET);
        d.box[4] = faces->type[MGP_BOTTOM] == MGP_DIRICHLET;
        d.box[5] = n[2] - (faces->type[MGP_TOP] == MGP_DIRICHLET);
        d.points = n[0] * n[1] * n[2];
        // r outside the box is read as stored (zero, as a fresh solver has it)
        if (!c.ok(mgp_fill_f64(&d.r, 0.0, stream))) return c.rc;
        c.L.push_back(d);
    }
    for (int64_t k = 0; k < ncycles && c.rc == MGP_OK; k++) {
        if (fc) c.fcycle(0);
        else c.vcycle(0);
    }
    Level& d0 = c.L[0];
    if (c.rc == MGP_OK && d0.u.data != u->data) {
        // the finest level ended in the second buffer: the result to the caller's u
        c.ok(mgp_memcpy_d2d(u->data, d0.u.data, (size_t)elements(d0.u) * sizeof(double), stream));
        Cycle::swap(d0);
    }
    if (c.rc == MGP_OK && norm) {
        c.ok(mgp_zero_scalars(norm, 1, stream));
        c.refresh(0);
        if (c.rc == MGP_OK &&
            mgp_residual_norm_f64(MGP_MODE_CONSISTENT, &d0.u, &d0.f, &d0.r, d0.box, d0.F.h, norm, stream) != MGP_OK)
            c.ok(mgp_residual_f64(MGP_MODE_CONSISTENT, &d0.u, &d0.f, &d0.r, d0.box, d0.F.h, norm, stream));
    }
    return c.rc;
}
