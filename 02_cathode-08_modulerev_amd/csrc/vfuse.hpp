// vfuse.hpp -- arguments of the BiCGStab work fused into the finest-level
// V-cycle kernels (vcycle.hpp), shared with the solver (mg.hpp).
#pragma once

namespace cathode {

// BiCGStab work fused into the finest-level kernels (real types; all other
// levels and the complex path pass pro = epi = 0):
//   down kernel, pro = 1: the V-cycle input is p = r (first) or
//       r + beta (p_old - om v) with p_old read from pin; pro = 2:
//       s = r - al v.  The new p / s is written to ps (never the buffer read)
//       and the (scaled, for FP32 levels) input to b of the level.
//   up kernel, epi = 1: y = V-cycle output (FP64), out = A y with the FP64
//       finest operator, block partials of <dotvec, out> (and <out, out>
//       with want_tt) into part[tile] and part[ntiles + tile].  Needs a halo
//       of 5 so the ring next to the tile is still valid after 4 half sweeps.
// Scalars come from sc at the slot indices given (beta = (rho1/rho)(al/om);
// the FP32 input scale is s32 * scale, the output scale 1 / scale).
// With pred set, the down kernel first sums the npred block partials at
// pred itself (every workgroup, same order) instead of reading a finished
// scalar: pro = 1 takes rho1 = (r0, r) from them (stored to sc[s_rho1] by
// workgroup 0; s_rho then names the previous iteration's slot), pro = 2
// takes (r0, v) and al = rho1 / (r0, v) (stored to sc[s_r0v], sc[s_al]).
// Warm start of a solve: x0 = sum_q a_q y_q over up to WARM_MAX vectors
// (each with its own row pitch, zero ghosts), (a_q) minimising |b - A x0|
// over their span.  report (pinned host memory, optional): the residual
// ratio |b - A x0|^2 / |b - A x0'|^2, x0' the optimum without the last
// vector, i.e. what the last vector adds.
constexpr int WARM_MAX = 4;
// dot products: <Ay_q, b> at q, <Ay_p, Ay_q> (p <= q) at WARM_MAX + warm_gidx(p, q), <b, b> last
constexpr int WARM_NP = WARM_MAX + WARM_MAX * (WARM_MAX + 1) / 2 + 1;
__host__ __device__ constexpr int warm_gidx(int p, int q) { return p * WARM_MAX - p * (p - 1) / 2 + (q - p); }
struct WarmSet {
    int k = 0;
    const double* y[WARM_MAX] = {nullptr, nullptr, nullptr, nullptr};
    int ld[WARM_MAX] = {0, 0, 0, 0};
    double* report = nullptr;
};

struct VFuse {
    int pro = 0, epi = 0, first = 0, mixed = 0, want_tt = 0;
    int s_s32 = 0, s_scale = 0, s_rho1 = 0, s_rho = 0, s_al = 0, s_om = 0, s_r0v = 0;
    const double *r = nullptr, *v = nullptr, *pin = nullptr;
    double* ps = nullptr;
    double* sc = nullptr;
    const double* pred = nullptr;
    int npred = 0;
    const double *aP = nullptr, *aE = nullptr, *aW = nullptr, *aN = nullptr, *aS = nullptr;
    double *y = nullptr, *out = nullptr;
    const double* dotvec = nullptr;
    double* part = nullptr;
};

}  // namespace cathode
