// FP16 storage of the coarse levels of an FP32 cycle (addition; the
// reference's hierarchy has one type T throughout).
//
// Fp16Levels keeps levels L .. L+m-1 below the last level L-1 of a
// PoissonSolver<float> as IEEE binary16 arrays (u, f, r per level, no halos)
// and runs the cycle's work on them with the FP16-storage kernels of
// include/mgp.h: red-black Gauss-Seidel half-sweeps (the coarsest level's
// passes in one LDS launch where it fits), residual, full weighting and
// trilinear prolongation with correction, in the order Vcycle / Fcycle run
// them on FP32 levels (vcycle.h:14-98, fcycle.h:13-97 of the reference).
// Arithmetic is FP32; only storage is FP16.
//
// Scaling: each restriction into an FP16 level multiplies by sigma = 2^-e,
// e the binary exponent of the max |r| of the level above (one device double
// per level, written by that level's residual kernel), and the prolongation
// back divides by it, so the values an FP16 level stores stay near 1 while
// the refinement drives the residual towards the FP64 rounding floor.
//
// Supported: consistent mode, Gauss-Seidel, full weighting, all faces
// Dirichlet (homogeneous: the correction equation of iterative refinement).
#ifndef POISSON_FP16LEVELS
#define POISSON_FP16LEVELS

#include <stdexcept>
#include <vector>

#include "domain.h"
#include "residual.h"

namespace Poisson {
class Fp16Levels : public CoarseLevels<float> {
public:
    // levels below `last` (the last FP32 level): count of them, coarsest-level
    // passes per visit (0 = nsmooth)
    Fp16Levels(const Domain<float>& last, uint_t count, int_t coarse_sweeps) : coarse_sweeps(coarse_sweeps) {
        if (count < 1) throw std::invalid_argument("Fp16Levels: no levels");
        int64_t n[3] = {last.u->shape[0], last.u->shape[1], last.u->shape[2]};
        double h = last.settings.h;
        void* p = nullptr;
        mgpx::check(mgpx::dev_alloc(&p, (count + 1) * sizeof(double)), "Fp16Levels");
        norms = static_cast<double*>(p);
        try {
            for (uint_t l = 0; l < count; l++) {
                Level v;
                for (int d = 0; d < 3; d++) {
                    if ((n[d] - 1) % 2 != 0 || n[d] < 5)
                        throw std::invalid_argument("Fp16Levels: a level shape does not coarsen 2:1");
                    n[d] = (n[d] - 1) / 2 + 1;
                }
                h *= 2.0;
                v.h = h;
                for (int d = 0; d < 3; d++) v.n[d] = n[d];
                lv.push_back(v);
                Level& w = lv.back();
                w.u = alloc(w);
                w.f = alloc(w);
                w.r = alloc(w);
            }
        } catch (...) {
            release();
            throw;
        }
        mgpx::check(mgp_zero_scalars(norms, count + 1, mgpx::stream()), "Fp16Levels");
        for (Level& v : lv) {
            const mgp_array u = desc(v, v.u), f = desc(v, v.f), r = desc(v, v.r);
            mgpx::check(mgp_fill_f16(&u, 0.0, mgpx::stream()), "Fp16Levels");
            mgpx::check(mgp_fill_f16(&f, 0.0, mgpx::stream()), "Fp16Levels");
            mgpx::check(mgp_fill_f16(&r, 0.0, mgpx::stream()), "Fp16Levels");
        }
        mgpx::check(mgp_stream_sync(mgpx::stream()), "Fp16Levels");
    }
    Fp16Levels(const Fp16Levels&) = delete;
    Fp16Levels& operator=(const Fp16Levels&) = delete;
    ~Fp16Levels() override { release(); }

    uint_t levels() const { return (uint_t)lv.size(); }
    // device views of level t's arrays ('u', 'f', 'r'), for inspection
    mgp_array array(uint_t t, char which) const {
        const Level& v = lv.at(t);
        return desc(v, which == 'u' ? v.u : (which == 'f' ? v.f : v.r));
    }

    // residual of the last FP32 level with its max |r| (norm slot 0), then
    // f_0 = sigma_0 FW(r) and u_0 = 0
    void enter(Domain<float>& d) override {
        const void* st = mgpx::stream();
        mgpx::check(mgp_zero_scalars(norms, 1, (void*)st), "fp16 levels: norm");
        residual_norm<float>(d, norms);
        const mgp_array r = d.r->desc(), f0 = array(0, 'f'), u0 = array(0, 'u');
        mgpx::check(mgp_restrict_fw_scaled_f32_f16(&r, &f0, norms, (void*)st), "fp16 levels: restriction");
        mgpx::check(mgp_fill_f16(&u0, 0.0, (void*)st), "fp16 levels: zero guess");
    }
    void leave(Domain<float>& d) override {
        const mgp_array u0 = array(0, 'u'), u = d.u->desc();
        mgpx::check(mgp_prolong_add_scaled_f16_f32(&u0, &u, norms, mgpx::stream()), "fp16 levels: prolongation");
    }
    void vcycle(float omega, int_t nsmooth) override { vcycle_at(0, omega, nsmooth); }
    void fcycle(float omega, int_t nsmooth) override {
        if (levels() >= 2) fcycle_at(0, omega, nsmooth);
        vcycle_at(0, omega, nsmooth);
    }

private:
    struct Level {
        int64_t n[3];
        double h;
        void* u = nullptr;
        void* f = nullptr;
        void* r = nullptr;
    };
    std::vector<Level> lv;
    double* norms = nullptr;  // slot t: max |r| of the level above level t
    int_t coarse_sweeps;

    static void* alloc(const Level& v) {
        void* p = nullptr;
        mgpx::check(mgpx::dev_alloc(&p, (size_t)(v.n[0] * v.n[1] * v.n[2]) * 2), "Fp16Levels");
        return p;
    }
    static mgp_array desc(const Level& v, void* data) {
        mgp_array a{};
        a.data = data;
        for (int d = 0; d < 3; d++) a.shape[d] = v.n[d];
        a.stride[2] = 1;
        a.stride[1] = v.n[2];
        a.stride[0] = v.n[1] * v.n[2];
        return a;
    }
    static void box(const Level& v, int64_t b[6]) {
        for (int d = 0; d < 3; d++) {
            b[2 * d] = 1;
            b[2 * d + 1] = v.n[d] - 1;
        }
    }
    void release() {
        for (Level& v : lv) {
            mgpx::dev_free(v.u);
            mgpx::dev_free(v.f);
            mgpx::dev_free(v.r);
        }
        lv.clear();
        if (norms) mgpx::dev_free(norms);
        norms = nullptr;
    }

    // nsmooth red-black sweeps (relaxation.h:19-94 with the Gauss-Seidel
    // smoothing multiplier 2: red then black, nsmooth times)
    void relax(uint_t t, float omega, int_t nsmooth) {
        const Level& v = lv[t];
        int64_t b[6];
        box(v, b);
        const mgp_array u = desc(v, v.u), f = desc(v, v.f);
        void* st = mgpx::stream();
        if (mgp_rbgs_sweeps_f16(&u, &f, b, omega, v.h, nsmooth, st) == MGP_OK) return;
        for (int_t s = 0; s < nsmooth; s++)
            for (int colour = 0; colour < 2; colour++)
                mgpx::check(mgp_rbgs_half_f16(&u, &f, b, omega, v.h, colour, st), "fp16 levels: smoother");
    }
    // residual of level t with its max |r| (slot t+1), f_{t+1} = sigma FW(r),
    // u_{t+1} = 0
    void descend(uint_t t) {
        const Level &v = lv[t], &c = lv[t + 1];
        int64_t b[6];
        box(v, b);
        const mgp_array u = desc(v, v.u), f = desc(v, v.f), r = desc(v, v.r);
        const mgp_array cf = desc(c, c.f), cu = desc(c, c.u);
        void* st = mgpx::stream();
        double* nrm = norms + t + 1;
        mgpx::check(mgp_zero_scalars(nrm, 1, st), "fp16 levels: norm");
        mgpx::check(mgp_residual_f16(&u, &f, &r, b, v.h, nrm, st), "fp16 levels: residual");
        mgpx::check(mgp_restrict_fw_scaled_f16(&r, &cf, nrm, st), "fp16 levels: restriction");
        mgpx::check(mgp_fill_f16(&cu, 0.0, st), "fp16 levels: zero guess");
    }
    void ascend(uint_t t) {
        const Level &v = lv[t], &c = lv[t + 1];
        const mgp_array cu = desc(c, c.u), u = desc(v, v.u);
        mgpx::check(mgp_prolong_add_scaled_f16(&cu, &u, norms + t + 1, mgpx::stream()), "fp16 levels: prolongation");
    }
    void vcycle_at(uint_t t, float omega, int_t nsmooth) {
        if (t + 1 >= levels()) {
            relax(t, omega, coarse_sweeps > 0 ? coarse_sweeps : nsmooth);
            return;
        }
        relax(t, omega, nsmooth);
        descend(t);
        vcycle_at(t + 1, omega, nsmooth);
        ascend(t);
        relax(t, omega, nsmooth);
    }
    void fcycle_at(uint_t t, float omega, int_t nsmooth) {
        relax(t, omega, nsmooth);
        descend(t);
        if (t + 2 < levels()) fcycle_at(t + 1, omega, nsmooth);
        vcycle_at(t + 1, omega, nsmooth);
        ascend(t);
    }
};
}  // namespace Poisson
#endif
