// neutral.hip -- ground-state neutral and metastable density updates on the
// NDD handle's grid and plasma geometry (FluidModule::update_rhon,
// include/fluidModule.hpp:4963-5411, and update_rhom, :5416-5792).
//
// Each update is one implicit r-weighted diffusion system on the plasma
// cells, assembled straight into the multigrid's level-0 arrays (with the
// multigrid's own level-0 preparation folded in, as the NDD assemblies do),
// one solve from x = 0, and one pass that writes the new density and every
// face flux.  Face roles come from the geometry's face codes: inner faces
// diffuse, wall faces carry the reference's explicit (neutrals) or implicit
// (metastables) wall flux, the open face the old outward flux, the axis
// nothing; an r face coded GF_INLET (the inlet segment, the reference's x1
// wall) feeds the gas where the cell centre lies at x <= x_in.
//
// The flux pass owns, per plasma cell, its low x and low r faces (inner or
// wall) and its high faces when they are boundaries, so every face has one
// writer; the open face's old value below it is read by the same thread
// before that thread overwrites it.  All expressions follow the source
// order of the restated loops, without multiply-add contraction.
#include <chrono>
#include <cmath>
#include <cstring>

#include "ndd_core.hpp"

namespace cathode {

enum { NP_N = 0, NP_M, NP_NOLD, NP_MOLD, NP_RHOE, NP_EXCSTEP, NP_SUPER, NP_EXCMETA, NP_EXC, NP_IONZ, NP_IONZSTEP,
       NP_GIX, NP_GIR, NP_NX, NP_NR, NP_NXW, NP_NRW, NP_MX, NP_MR, NP_MXW, NP_MRW, NP_ZERO, NP_COUNT };

enum { SPN = 0, SPM = 1 };  // neutrals, metastables

struct NeutArgs {
    const double* n_old;  // the species' density at the start of the update
    const double* nD;     // the density whose face means set D (neutrals: n_old; metastables: rhon)
    const double* rhoe;
    const double *g0, *g1;      // neutrals: nu_excStep, nu_super; metastables: nu_excMeta, nu_exc
    const double *l0, *l1, *l2; // neutrals: nu_excMeta, nu_exc, nu_ionz; metastables: nu_ionzStep, nu_excStep, nu_super
    const double *Gix, *Gir, *Gmx, *Gmr;  // wall fluxes the neutral walls return
    double *fx, *fr, *fxw, *frw;          // the species' fluxes and wall fluxes (in: old, out: new)
    double* n;                            // the species' density (out)
    const double* sol;                    // the solve's solution, pitch ld0
    int* neg;                             // count of cells reset from a negative value
    double DN, dt, vth, fn_In, x_in, xL;
    int guard;                            // metastables: blocks 1..guard reset negative values
};

// cell-centre x of column i (initialModule.hpp:84)
__device__ __forceinline__ double cell_x(const NddView& V, const NeutArgs& A, int i)
{
#pragma clang fp contract(off)
    return A.xL + ((double)i - 0.5) * V.dx;
}

// one row of the neutral (fluidModule.hpp:5001-5112) or metastable
// (:5449-5573) system
template <class G>
__device__ __forceinline__ Row5 neut_row(const G& g, const NddView& V, const NeutArgs& A, int sp, int i, int j)
{
#pragma clang fp contract(off)
    const double dx = V.dx, dr = V.dr;
    const double* r = V.r;
    const size_t k = IX(i, j);
    const double rL = (r[j] + r[j - 1]) / 2.0;
    const double rR = (r[j] + r[j + 1]) / 2.0;
    const double qL = (r[j] + r[j - 1]) / 2.0 / r[j];
    const double qR = (r[j] + r[j + 1]) / 2.0 / r[j];
    const double nc = A.nD[k];
    const double D_Lx = A.DN / ((nc + A.nD[IX(i - 1, j)]) / 2.0 + 1e-100);
    const double D_Rx = A.DN / ((nc + A.nD[IX(i + 1, j)]) / 2.0 + 1e-100);
    const double D_Lr = A.DN / ((nc + A.nD[IX(i, j - 1)]) / 2.0 + 1e-100);
    const double D_Rr = A.DN / ((nc + A.nD[IX(i, j + 1)]) / 2.0 + 1e-100);
    const int tL = g.xf(V, i, j) & GF_TYPE, tR = g.xf(V, i + 1, j) & GF_TYPE;
    const int tS = g.rf(V, i, j) & GF_TYPE, cN = g.rf(V, i, j + 1), tN = cN & GF_TYPE;
    const double bLx = (double)(tL == GF_INNER), bRx = (double)(tR == GF_INNER);
    const double bLr = (double)(tS == GF_INNER), bRr = (double)(tN == GF_INNER);
    Row5 R;
    R.E = D_Rx * dr * dr * bRx;
    R.W = D_Lx * dr * dr * bLx;
    R.N = qR * D_Rr * dx * dx * bRr;
    R.S = qL * D_Lr * dx * dx * bLr;
    const double diff = (D_Rx * bRx + D_Lx * bLx) * dr * dr + (qR * D_Rr * bRr + qL * D_Lr * bLr) * dx * dx;
    const double ne = A.rhoe[k];
    if (sp == SPN) {
        // walls return the ion and metastable wall fluxes; the inlet adds the feed
        double GwLx = 0.0, GwRx = 0.0, GwLr = 0.0, GwRr = 0.0, Gin = 0.0, Gop = 0.0;
        if (tL == GF_WLO) GwLx = -A.Gix[k] - A.Gmx[k];
        if (tR == GF_WHI) GwRx = -A.Gix[IX(i + 1, j)] - A.Gmx[IX(i + 1, j)];
        if (tS == GF_WLO) GwLr = -A.Gir[k] - A.Gmr[k];
        if (tN == GF_WHI) {
            const size_t kN = IX(i, j + 1);
            if ((cN & GF_INLET) && cell_x(V, A, i) <= A.x_in)
                Gin = -A.Gir[kN] - A.Gmr[kN] - A.fn_In;
            else
                GwRr = -A.Gir[kN] - A.Gmr[kN];
        }
        if (tN == GF_OPEN) Gop = fmax(rL / rR * A.fr[k], 0.0);
        R.P = diff + dx * dx * dr * dr / A.dt;
        const double Gs = (A.g0[k] + A.g1[k]) * ne;
        const double Ls = (A.l0[k] + 0.5 * A.l1[k] + A.l2[k]) * ne;
        R.b = A.n_old[k] / A.dt * dx * dx * dr * dr + (Gs - Ls) * dx * dx * dr * dr + (GwLx) * dr * dx * dr -
              (GwRx) * dr * dx * dr + qL * (GwLr) * dx * dx * dr - qR * (GwRr + Gop + Gin) * dx * dx * dr;
    } else {
        // thermal wall loss per density, implicit; every wall (the inlet one too)
        const double gLx = tL == GF_WLO ? -0.25 * A.vth : 0.0;
        const double gRx = tR == GF_WHI ? 0.25 * A.vth : 0.0;
        const double gLr = tS == GF_WLO ? -0.25 * A.vth : 0.0;
        const double gRr = tN == GF_WHI ? 0.25 * A.vth : 0.0;
        const double Gop = tN == GF_OPEN ? rL / rR * A.fr[k] : 0.0;
        R.P = diff - gLx * dr * dx * dr + gRx * dr * dx * dr - qL * gLr * dx * dx * dr + qR * gRr * dx * dx * dr +
              dx * dx * dr * dr / A.dt;
        const double Gs = ne * (A.g0[k] + 0.5 * A.g1[k]);
        const double Ls = ne * (A.l0[k] + A.l1[k] + A.l2[k]);
        // (the reference's explicit wall terms of this row are identically 0.0)
        R.b = A.n_old[k] / A.dt * dx * dx * dr * dr + (Gs - Ls) * dx * dx * dr * dr - qR * (Gop) * dx * dx * dr;
    }
    return R;
}

template <class G>
__global__ void k_neut_assemble(G g, NddView V, NeutArgs A, int sp, double* aP, double* aE, double* aW, double* aN,
                                double* aS, double* b, double* s_out, double* part)
{
    THREAD_IJ(1, 1);
    double amax = 0.0, bb = 0.0;
    if (i <= V.ni && j <= V.nj) {
        if (g.act(V, i, j)) {
            Row5 R = neut_row(g, V, A, sp, i, j);
            prep_row(g, V, i, j, R, s_out, amax, bb);
            store_row(V, i, j, R, aP, aE, aW, aN, aS, b);
        } else {
            store_idle_row(V, i, j, aP, aE, aW, aN, aS, b, s_out);
        }
    }
    prep_partials(amax, bb, part);
}

// the new density of plasma cell (i, j): the solution, where a negative
// metastable value in a guard block is put back to max(old, 0) before the
// fluxes are taken (fluidModule.hpp:5582-5595); neutrals take their fluxes
// from the solution itself and are reset afterwards (:5383-5399)
template <class G>
__device__ __forceinline__ double new_density(const G& g, const NddView& V, const NeutArgs& A, int sp, int i, int j)
{
    const double s = A.sol[IX0(i, j)];
    if (sp == SPM && s < 0.0 && g.blk(V, i, j) <= A.guard) return fmax(A.n_old[IX(i, j)], 0.0);
    return s;
}

// the face-flux update (neutrals :5130-5262, metastables :5597-5693) and
// the new density
template <class G>
__global__ void k_neut_flux(G g, NddView V, NeutArgs A, int sp)
{
#pragma clang fp contract(off)
    THREAD_IJ(1, 1);
    if (i > V.ni || j > V.nj || !g.act(V, i, j)) return;
    const double dx = V.dx, dr = V.dr;
    const double* r = V.r;
    const size_t k = IX(i, j), kE = IX(i + 1, j), kN = IX(i, j + 1);
    const int tL = g.xf(V, i, j) & GF_TYPE, tR = g.xf(V, i + 1, j) & GF_TYPE;
    const int tS = g.rf(V, i, j) & GF_TYPE, cN = g.rf(V, i, j + 1), tN = cN & GF_TYPE;
    const double s = A.sol[IX0(i, j)];
    const double nc = new_density(g, V, A, sp, i, j);
    const double f_below = A.fr[k];  // the old flux of this cell's low r face
    const double nD = A.nD[k];
    // low x face
    if (tL == GF_INNER) {
        const double D = A.DN / ((nD + A.nD[IX(i - 1, j)]) / 2.0 + 1e-100);
        A.fx[k] = -D * (nc - new_density(g, V, A, sp, i - 1, j)) / dx;
    } else if (tL == GF_WLO) {
        A.fx[k] = sp == SPN ? -A.Gix[k] - A.Gmx[k] : -0.25 * nc * A.vth;
        A.fxw[k] = A.fx[k];
    }
    // high x face
    if (tR == GF_WHI) {
        A.fx[kE] = sp == SPN ? -A.Gix[kE] - A.Gmx[kE] : 0.25 * nc * A.vth;
        A.fxw[kE] = A.fx[kE];
    }
    // low r face (the axis carries nothing)
    if (tS == GF_INNER) {
        const double D = A.DN / ((nD + A.nD[IX(i, j - 1)]) / 2.0 + 1e-100);
        A.fr[k] = -D * (nc - new_density(g, V, A, sp, i, j - 1)) / dr;
    } else if (tS == GF_WLO) {
        A.fr[k] = sp == SPN ? -A.Gir[k] - A.Gmr[k] : -0.25 * nc * A.vth;
        A.frw[k] = A.fr[k];
    }
    // high r face
    if (tN == GF_WHI) {
        if (sp == SPN && (cN & GF_INLET) && cell_x(V, A, i) < A.x_in) {
            A.fr[kN] = -A.Gir[kN] - A.Gmr[kN] - A.fn_In;  // the inlet keeps its wall flux
        } else {
            A.fr[kN] = sp == SPN ? -A.Gir[kN] - A.Gmr[kN] : 0.25 * nc * A.vth;
            A.frw[kN] = A.fr[kN];
        }
    } else if (tN == GF_OPEN) {
        const double rL = (r[j] + r[j - 1]) / 2.0;
        const double rR = (r[j] + r[j + 1]) / 2.0;
        A.fr[kN] = sp == SPN ? fmax(rL / rR * f_below, 0.0) : f_below * rL / rR;
        A.frw[kN] = 0.0;
    }
    // the density: neutrals reset every negative value now, metastables took it above
    double out = nc;
    bool reset = false;
    if (sp == SPN) {
        reset = s < 0.0;
        if (reset) out = fmax(A.n_old[k], 0.0);
    } else {
        reset = s < 0.0 && g.blk(V, i, j) <= A.guard;
    }
    A.n[k] = out;
    if (reset) atomicAdd(A.neg, 1);
}

}  // namespace cathode

using namespace cathode;

namespace {

bool bad_number(double v) { return !std::isfinite(v); }

}  // namespace

// the update on host arrays (dev = false: staged through the handle's
// planes) or on the caller's device arrays (dev = true)
static int neutral_entry(void* handle, const cathode_neutral_params* p, cathode_neutral_fields* f,
                         cathode_neutral_report* rep, bool dev)
{
    try {
        NddHandle* H = (NddHandle*)handle;
        if (!H || !p || !f) {
            set_error("ndd_neutral_update: null argument");
            return CATHODE_ERR_ARG;
        }
        const bool do_n = (p->which & 1) != 0, do_m = (p->which & 2) != 0;
        if ((p->which & ~3) || !(do_n || do_m) || (p->mode != CATHODE_MODE_FP64 && p->mode != CATHODE_MODE_MIXED) ||
            p->maxITR < 1 || !(p->tol > 0.0) || p->flag_error < 0 || p->flag_error > 3 ||
            (do_n && (!(p->dt_n > 0.0) || !(p->Tn > 0.0) || !(p->ri > 0.0) || !(p->massi > 0.0) ||
                      bad_number(p->fn_In) || bad_number(p->x_in) || bad_number(p->xL))) ||
            (do_m && (!(p->dt_m > 0.0) || !(p->Tn > 0.0) || !(p->massi > 0.0) || bad_number(p->DmN) ||
                      p->guard_blocks < 0))) {
            set_error("ndd_neutral_update: invalid parameters");
            return CATHODE_ERR_ARG;
        }
        if (!f->rhon || (do_n && (!f->rhoUnx || !f->rhoUnr || !f->rhoUnx_wall || !f->rhoUnr_wall)) ||
            (do_m && (!f->rhom || !f->rhoUmx || !f->rhoUmr || !f->rhoUmx_wall || !f->rhoUmr_wall))) {
            set_error("ndd_neutral_update: a density or flux array of an updated species is missing");
            return CATHODE_ERR_ARG;
        }
        auto t0 = std::chrono::steady_clock::now();
        const NddView& V = H->V;
        const size_t n = H->n;
        H->neut.alloc(NP_COUNT * n);
        H->neut_cnt.alloc(2);
        double* P[NP_COUNT];
        for (int q = 0; q < NP_COUNT; q++) P[q] = H->neut.p + q * n;
        CATHODE_HIP(hipMemsetAsync(P[NP_ZERO], 0, n * sizeof(double), H->s));
        CATHODE_HIP(hipMemsetAsync(H->neut_cnt.p, 0, 2 * sizeof(int), H->s));
        // host -> device, or the caller's device arrays; an absent input
        // reads as the zero plane
        const double* zero = P[NP_ZERO];
        auto up = [&](int q, const double* src) -> const double* {
            if (!src) return zero;
            if (dev) {
                P[q] = const_cast<double*>(src);
                return src;
            }
            CATHODE_HIP(hipMemcpyAsync(P[q], src, n * sizeof(double), hipMemcpyHostToDevice, H->s));
            return P[q];
        };
        up(NP_N, f->rhon);
        if (do_m) up(NP_M, f->rhom);
        const double* rhoe = up(NP_RHOE, f->rhoe);
        const double* excStep = up(NP_EXCSTEP, f->nu_excStep);
        const double* super = up(NP_SUPER, f->nu_super);
        const double* excMeta = up(NP_EXCMETA, f->nu_excMeta);
        const double* exc = up(NP_EXC, f->nu_exc);
        const double* ionz = do_n ? up(NP_IONZ, f->nu_ionz) : zero;
        const double* ionzStep = do_m ? up(NP_IONZSTEP, f->nu_ionzStep) : zero;
        const double* Gix = do_n ? up(NP_GIX, f->rhoUix_wall) : zero;
        const double* Gir = do_n ? up(NP_GIR, f->rhoUir_wall) : zero;
        if (do_n) {
            up(NP_NX, f->rhoUnx);
            up(NP_NR, f->rhoUnr);
            up(NP_NXW, f->rhoUnx_wall);
            up(NP_NRW, f->rhoUnr_wall);
        }
        // the metastable wall fluxes: read by the neutral walls, updated by the metastables
        const double* Gmx = up(NP_MXW, f->rhoUmx_wall);
        const double* Gmr = up(NP_MRW, f->rhoUmr_wall);
        if (do_m) {
            up(NP_MX, f->rhoUmx);
            up(NP_MR, f->rhoUmr);
        }

        constexpr double KB = K_BOLTZ;
        cathode_status st[2] = {};
        double err[2] = {0.0, 0.0};
        for (int sp = SPN; sp <= SPM; sp++) {
            if (sp == SPN ? !do_n : !do_m) continue;
            NeutArgs A{};
            const int pn = sp == SPN ? NP_N : NP_M, po = sp == SPN ? NP_NOLD : NP_MOLD;
            A.n_old = P[po];
            A.nD = sp == SPN ? P[NP_NOLD] : P[NP_N];
            A.rhoe = rhoe;
            if (sp == SPN) {
                A.g0 = excStep; A.g1 = super;
                A.l0 = excMeta; A.l1 = exc; A.l2 = ionz;
                const double sigma = M_PI * (2.0 * p->ri) * (2.0 * p->ri);  // fluidModule.hpp:4979-4981
                A.DN = 2.0 / (3.0 * sigma) * std::sqrt(KB * p->Tn / (M_PI * p->massi));
                A.dt = p->dt_n;
                A.fx = P[NP_NX]; A.fr = P[NP_NR]; A.fxw = P[NP_NXW]; A.frw = P[NP_NRW];
            } else {
                A.g0 = excMeta; A.g1 = exc;
                A.l0 = ionzStep; A.l1 = excStep; A.l2 = super;
                A.DN = p->DmN;
                A.dt = p->dt_m;
                A.fx = P[NP_MX]; A.fr = P[NP_MR]; A.fxw = P[NP_MXW]; A.frw = P[NP_MRW];
            }
            A.vth = std::sqrt(8.0 * KB * p->Tn / (M_PI * p->massi));  // :5433
            A.Gix = Gix; A.Gir = Gir; A.Gmx = Gmx; A.Gmr = Gmr;
            A.n = P[pn];
            A.neg = H->neut_cnt.p + sp;
            A.fn_In = p->fn_In; A.x_in = p->x_in; A.xL = p->xL;
            A.guard = p->guard_blocks;
            d2d(*H, {{P[po], P[pn]}});
            NDD_GEO_LAUNCH(*H, k_neut_assemble, ggrid(V.ni, V.nj), gblk(), V, A, sp, H->mg.aP(), H->mg.aE(),
                           H->mg.aW(), H->mg.aN(), H->mg.aS(), H->mg.rhs(), H->mg.prep_s(), H->mg.prep_part());
            CATHODE_LAUNCH_CHECK();
            H->mg.prepared = true;
            H->mg.same_operator = false;
            H->mg.keep_guess = false;
            H->mg.expected_its = 0;
            H->mg.side_cnv = -INFINITY;
            H->mg.solve(p->maxITR, p->tol, p->mode, H->s, &st[sp]);
            A.sol = H->mg.sol();
            NDD_GEO_LAUNCH(*H, k_neut_flux, ggrid(V.ni, V.nj), gblk(), V, A, sp);
            CATHODE_LAUNCH_CHECK();
            // calcRes over the plasma cells (the face pairs of the call are not used)
            double e3[3];
            const ResPairs RP{P[pn], P[po], P[pn], P[po], P[pn], P[po]};
            calcres3(*H, RP, e3, p->flag_error);
            err[sp] = e3[0];
        }
        // device -> host
        auto down = [&](double* dst, int q) {
            if (!dev) CATHODE_HIP(hipMemcpyAsync(dst, P[q], n * sizeof(double), hipMemcpyDeviceToHost, H->s));
        };
        if (do_n) {
            down(f->rhon, NP_N);
            down(f->rhoUnx, NP_NX);
            down(f->rhoUnr, NP_NR);
            down(f->rhoUnx_wall, NP_NXW);
            down(f->rhoUnr_wall, NP_NRW);
        }
        if (do_m) {
            down(f->rhom, NP_M);
            down(f->rhoUmx, NP_MX);
            down(f->rhoUmr, NP_MR);
            down(f->rhoUmx_wall, NP_MXW);
            down(f->rhoUmr_wall, NP_MRW);
        }
        int neg[2] = {0, 0};
        CATHODE_HIP(hipMemcpyAsync(neg, H->neut_cnt.p, 2 * sizeof(int), hipMemcpyDeviceToHost, H->s));
        CATHODE_HIP(hipStreamSynchronize(H->s));
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rep) {
            std::memset(rep, 0, sizeof(*rep));
            rep->solve_n = st[SPN];
            rep->solve_m = st[SPM];
            rep->solve_n.seconds = rep->solve_m.seconds = secs;
            rep->error_rhon = err[SPN];
            rep->error_rhom = err[SPM];
            rep->neg_rhon = neg[SPN];
            rep->neg_rhom = neg[SPM];
        }
        return CATHODE_OK;
    CATHODE_CATCH_END
}

extern "C" int cathode_ndd_neutral_update(void* handle, const cathode_neutral_params* p, cathode_neutral_fields* f,
                                          cathode_neutral_report* rep)
{
    return neutral_entry(handle, p, f, rep, false);
}

extern "C" int cathode_ndd_neutral_update_dev(void* handle, const cathode_neutral_params* p,
                                              cathode_neutral_fields* f, cathode_neutral_report* rep)
{
    return neutral_entry(handle, p, f, rep, true);
}
