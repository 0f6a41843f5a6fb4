// transport.hip -- transport coefficients from the rate tables on the NDD
// handle's grid and geometry (FluidModule::update_transport_coef,
// include/fluidModule.hpp:5799-5976).
//
// One pointwise pass over the extended grid after one max reduction of n_e
// over the plasma cells: on the plasma cells the electron drift's Mach number
// and the ion-acoustic anomalous collision frequency, then on every cell the
// collision frequencies (table lookups in Te times n_n or n_m), the ion-
// neutral frequency (lookup in m_i |U_i|^2), the energy-loss rate and its
// Te derivative, and the magnetised mobility tensor.  The tables (about 15 KB
// each) stay in global memory, where the two entries a cell reads per table
// come from L2; the pass is bound by its ~290 B per cell of HBM traffic.
// Expressions follow the restated loops without multiply-add contraction.
#include <cmath>
#include <cstring>
#include <vector>

#include "ndd_core.hpp"

namespace cathode {

enum { TB_ELAS = 0, TB_EXC, TB_RESO, TB_META, TB_IONZ, TB_IONZSTEP, TB_EXCSTEP, TB_SUPER, TB_ELOSS_N, TB_ELOSS_M,
       TB_COUNT };

struct TransportArgs {
    const double *rhoe, *rhoi, *rhon, *rhom, *Te, *Bx, *Br, *nu_ano, *Uix, *Uir, *Uip, *rhoUex, *rhoUer;
    double* iat;
    double* out[22];   // nu_elas .. nu_super (8), nu_m, nu_m1, rate_ionize, nui_m, rate_eloss, delossdeps,
                       // mu_para, mu_xx, mu_rr, mu_xr, mu_perp, mu_min, Mache (NULL: not written)
    const double* tab[TB_COUNT];
    const double* nui_tab;
    int n_e, n_i;
    double dTe, dEi, masse, massi, Ti, nu_eff, flag_Bohm, flag_Sagdeev;
    const double* rhoe_max;  // device scalar
};
enum { TO_M = 8, TO_M1, TO_RI, TO_NUI, TO_EL, TO_DEL, TO_MUP, TO_MXX, TO_MRR, TO_MXR, TO_MPERP, TO_MMIN, TO_MACH };

// miniFuncs.hpp:35-46; the reference compares iR with the table size as an
// unsigned value.  An index below 0 (Te <= -dTe, outside the reference's
// defined behaviour) reads entry 0.
__device__ __forceinline__ double tab_Te(const double* v, int n, double Te, double dTe)
{
#pragma clang fp contract(off)
    int iL = (int)(Te / dTe);
    const int iR = (int)(Te / dTe) + 1;
    const double TeL = iL * dTe;
    if ((unsigned)iR < (unsigned)(n - 1)) {
        iL = iL < 0 ? 0 : iL;
        return (v[iR] - v[iL]) / (dTe + 1e-100) * (Te - TeL) + v[iL];
    }
    return v[n - 1];
}

// miniFuncs.hpp:53-67
__device__ __forceinline__ double tab_dTe(const double* v, int n, double Te, double dTe)
{
#pragma clang fp contract(off)
    int iL = (int)(Te / dTe);
    int iR = (int)(Te / dTe) + 1;
    if (!((unsigned)iR < (unsigned)(n - 1))) {
        iL = n - 2;
        iR = n - 1;
    }
    const double TeL = iL * dTe;
    const double TeR = iR * dTe;
    return (v[iR] - v[iL < 0 ? 0 : iL]) / (TeR - TeL + 1e-100);
}

// miniFuncs.hpp:75-91
__device__ __forceinline__ double tab_Ei(const double* v, int n, double Ei, double dEi, double massi)
{
#pragma clang fp contract(off)
    int iL = (int)(Ei / dEi);
    const int iR = (int)(Ei / dEi) + 1;
    const double EiL = iL * dEi;
    if ((unsigned)iR < (unsigned)(n - 1)) {
        iL = iL < 0 ? 0 : iL;
        return (v[iR] - v[iL]) / (dEi + 1e-100) * (Ei - EiL) + v[iL];
    }
    const double imax = n - 1;
    const double v_max = sqrt(imax * dEi / massi);
    const double v_tmp = sqrt(Ei / massi);
    const double sigma_max = v[n - 1] / v_max;
    return sigma_max * v_tmp;
}

// max of n_e over the plasma cells, one partial per block
template <class G>
__global__ void k_rhoe_max_part(G g, NddView V, const double* __restrict__ rhoe, double* __restrict__ part)
{
    __shared__ double lds[NB_X * NB_Y / 64];
    THREAD_IJ(1, 1);
    double v = -INFINITY;
    if (i <= V.ni && j <= V.nj && g.act(V, i, j)) v = rhoe[IX(i, j)];
    const double m = block_max<NB_X * NB_Y>(v, lds);
    if (lin_tid() == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = m;
}

// the reference starts its maximum at 0
__global__ void k_rhoe_max_finish(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    __shared__ double lds[256 / 64];
    double v = -INFINITY;
    for (int k = threadIdx.x; k < nb; k += 256) v = fmax(v, part[k]);
    const double m = block_max<256>(v, lds);
    if (threadIdx.x == 0) out[0] = fmax(0.0, m);
}

template <class G>
__global__ void k_transport(G g, NddView V, TransportArgs A)
{
#pragma clang fp contract(off)
    THREAD_IJ(0, 0);
    if (i > V.ni + 1 || j > V.nj + 1) return;
    const double* r = V.r;
    const size_t k = IX(i, j);
    const double Te = A.Te[k];
    const bool interior = i >= 1 && i <= V.ni && j >= 1 && j <= V.nj;
    // :5826-5899, the plasma cells
    if (interior && g.act(V, i, j)) {
        const int tL = g.xf(V, i, j) & GF_TYPE, tR = g.xf(V, i + 1, j) & GF_TYPE;
        const int tS = g.rf(V, i, j) & GF_TYPE, tN = g.rf(V, i, j + 1) & GF_TYPE;
        const double bLx = (double)(tL != GF_WLO), bRx = (double)(tR != GF_WHI);
        const double bLr = (double)(tS != GF_WLO), bRr = (double)(tN != GF_WHI);
        const double ut = sqrt(2.0 * K_BOLTZ * Te / A.masse);
        const double rL = (r[j] + r[j - 1]) / 2.0;
        const double rR = (r[j] + r[j + 1]) / 2.0;
        const double fx = (A.rhoUex[IX(i + 1, j)] * bRx + A.rhoUex[k] * bLx) / 2.0;
        const double fr = (rR * A.rhoUer[IX(i, j + 1)] * bRr + rL * A.rhoUer[k] * bLr) / (2.0 * r[j]);
        const double rmax = A.rhoe_max[0];
        const double Uex = fx / (A.rhoe[k] + rmax * 0.001);
        const double Uer = fr / (A.rhoe[k] + rmax * 0.001);
        const double Uix = (A.Uix[k] + A.Uix[IX(i + 1, j)]) / 2.0;
        const double Uir = (rL * A.Uir[k] + rR * A.Uir[IX(i, j + 1)]) / (2.0 * r[j]);
        if (A.out[TO_MACH]) A.out[TO_MACH][k] = sqrt(Uex * Uex + Uer * Uer) / (ut + 100.0);
        const double cs = sqrt(2.0 * K_BOLTZ * (Te + A.Ti) / A.massi);
        const double omega_pi = sqrt(A.rhoi[k] * K_E0 * K_E0 / (A.massi * K_EPS0));
        const double dx_ = Uex - Uix, dr_ = Uer - Uir;
        const double ud = sqrt(dx_ * dx_ + dr_ * dr_);
        A.iat[k] = 0.01 * omega_pi * ud / cs * Te / A.Ti;
    }
    // :5901-5974, every cell
    const double nn = A.rhon[k], nm = A.rhom[k];
    const int ne = A.n_e;
    double nu[8];
#pragma unroll
    for (int q = 0; q < 8; q++) nu[q] = (q < TB_IONZSTEP ? nn : nm) * tab_Te(A.tab[q], ne, Te, A.dTe);
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (A.out[q]) A.out[q][k] = nu[q];
    const double nu_m_real = nn * tab_Te(A.tab[TB_ELAS], ne, Te, A.dTe);
    const double nu_m = nu_m_real + A.flag_Bohm * A.nu_ano[k] + A.flag_Sagdeev * A.iat[k];
    double vv = 0.0;
    if (interior) {
        const double rL = (r[j] + r[j - 1]) / 2.0;
        const double rR = (r[j] + r[j + 1]) / 2.0;
        const double Uix = (A.Uix[k] + A.Uix[IX(i + 1, j)]) / 2.0;
        const double Uir = (rL * A.Uir[k] + rR * A.Uir[IX(i, j + 1)]) / (2.0 * r[j]);
        const double Uip = A.Uip[k];
        vv = Uix * Uix + Uir * Uir + Uip * Uip;
    }
    const double Ei = A.massi * vv;
    const double c = 2.0 / (3.0 * K_BOLTZ);
    const double mu_para = K_E0 / (A.masse * nu_m + 1e-100);
    const double Bx = A.Bx[k], Br = A.Br[k];
    const double Bmag = sqrt(Bx * Bx + Br * Br);
    const double W = K_E0 * Bmag / (A.masse * nu_m + 1e-100);
    const double Wx = K_E0 * Bx / (A.masse * nu_m + 1e-100);
    const double Wr = K_E0 * Br / (A.masse * nu_m + 1e-100);
    const double deno = 1.0 + W * W;
    double* const* o = A.out;
    if (o[TO_M]) o[TO_M][k] = nu_m;
    if (o[TO_M1]) o[TO_M1][k] = nu_m_real + A.nu_eff;
    if (o[TO_RI]) o[TO_RI][k] = A.rhoe[k] * (nu[TB_IONZ] + nu[TB_IONZSTEP]);
    if (o[TO_NUI]) o[TO_NUI][k] = nn * tab_Ei(A.nui_tab, A.n_i, Ei, A.dEi, A.massi);
    if (o[TO_EL])
        o[TO_EL][k] = nn * tab_Te(A.tab[TB_ELOSS_N], ne, Te, A.dTe) + nm * tab_Te(A.tab[TB_ELOSS_M], ne, Te, A.dTe);
    if (o[TO_DEL])
        o[TO_DEL][k] = c * nn * tab_dTe(A.tab[TB_ELOSS_N], ne, Te, A.dTe) + c * nm * tab_dTe(A.tab[TB_ELOSS_M], ne, Te, A.dTe);
    if (o[TO_MUP]) o[TO_MUP][k] = mu_para;
    if (o[TO_MXX]) o[TO_MXX][k] = mu_para * (1.0 + Wx * Wx) / deno;
    if (o[TO_MRR]) o[TO_MRR][k] = mu_para * (1.0 + Wr * Wr) / deno;
    if (o[TO_MXR]) o[TO_MXR][k] = mu_para * (Wx * Wr) / deno;
    if (o[TO_MPERP]) o[TO_MPERP][k] = mu_para / deno;
    if (o[TO_MMIN]) o[TO_MMIN][k] = mu_para * 2.0 / deno;
}

}  // namespace cathode

using namespace cathode;

extern "C" int cathode_ndd_set_rate_tables(void* handle, const cathode_rate_tables* t)
{
    try {
        NddHandle* H = (NddHandle*)handle;
        if (!H || !t || t->n_e < 2 || t->n_i < 2 || !(t->dTe > 0.0) || !(t->dEi > 0.0) || !t->nui_m) {
            set_error("ndd_set_rate_tables: need at least two entries per table and positive steps");
            return CATHODE_ERR_ARG;
        }
        const double* src[TB_COUNT] = {t->nu_elas, t->nu_exc, t->nu_excReso, t->nu_excMeta, t->nu_ionz,
                                       t->nu_ionzStep, t->nu_excStep, t->nu_super, t->eloss_n, t->eloss_m};
        for (const double* p : src)
            if (!p) {
                set_error("ndd_set_rate_tables: a table is missing");
                return CATHODE_ERR_ARG;
            }
        const size_t ne = (size_t)t->n_e, ni = (size_t)t->n_i;
        std::vector<double> all(TB_COUNT * ne + ni);
        for (int q = 0; q < TB_COUNT; q++) std::memcpy(all.data() + q * ne, src[q], ne * sizeof(double));
        std::memcpy(all.data() + TB_COUNT * ne, t->nui_m, ni * sizeof(double));
        H->rate_tab.alloc(all.size());
        CATHODE_HIP(hipMemcpyAsync(H->rate_tab.p, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice, H->s));
        CATHODE_HIP(hipStreamSynchronize(H->s));
        H->rate_n[0] = t->n_e;
        H->rate_n[1] = t->n_i;
        H->rate_d[0] = t->dTe;
        H->rate_d[1] = t->dEi;
        return CATHODE_OK;
    CATHODE_CATCH_END
}

static int transport_entry(void* handle, const cathode_transport_params* p, cathode_transport_fields* f, bool dev)
{
    try {
        NddHandle* H = (NddHandle*)handle;
        if (!H || !p || !f) {
            set_error("ndd_transport_update: null argument");
            return CATHODE_ERR_ARG;
        }
        if (H->rate_n[0] < 2) {
            set_error("ndd_transport_update: no rate tables (cathode_ndd_set_rate_tables)");
            return CATHODE_ERR_ARG;
        }
        if (!(p->masse > 0.0) || !(p->massi > 0.0) || !(p->Ti > 0.0) || !std::isfinite(p->nu_eff) ||
            !f->Te || !f->nu_ano_IAT) {
            set_error("ndd_transport_update: invalid parameters, or Te / nu_ano_IAT missing");
            return CATHODE_ERR_ARG;
        }
        const NddView& V = H->V;
        const size_t n = H->n;
        constexpr int NIN = 13, NOUT = 22;
        const double* in_src[NIN] = {f->rhoe, f->rhoi, f->rhon, f->rhom, f->Te, f->Bx, f->Br, f->nu_ano,
                                     f->Uix, f->Uir, f->Uip, f->rhoUex, f->rhoUer};
        double* out_dst[NOUT] = {f->nu_elas, f->nu_exc, f->nu_excReso, f->nu_excMeta, f->nu_ionz, f->nu_ionzStep,
                                 f->nu_excStep, f->nu_super, f->nu_m, f->nu_m1, f->rate_ionize, f->nui_m,
                                 f->rate_eloss, f->delossdeps, f->mu_para, f->mu_xx, f->mu_rr, f->mu_xr,
                                 f->mu_perp, f->mu_min, f->Mache, nullptr};
        // planes: inputs, nu_ano_IAT, outputs, zero
        const int NPL = NIN + 1 + NOUT + 1;
        double* P[NPL] = {};
        H->trp.alloc(dev ? 1 * n : NPL * n);
        double* zero = dev ? H->trp.p : H->trp.p + (NPL - 1) * n;
        CATHODE_HIP(hipMemsetAsync(zero, 0, n * sizeof(double), H->s));
        TransportArgs A{};
        const double** in_dst[NIN] = {&A.rhoe, &A.rhoi, &A.rhon, &A.rhom, &A.Te, &A.Bx, &A.Br, &A.nu_ano,
                                      &A.Uix, &A.Uir, &A.Uip, &A.rhoUex, &A.rhoUer};
        for (int q = 0; q < NIN; q++) {
            if (!in_src[q]) {
                *in_dst[q] = zero;
            } else if (dev) {
                *in_dst[q] = in_src[q];
            } else {
                P[q] = H->trp.p + q * n;
                CATHODE_HIP(hipMemcpyAsync(P[q], in_src[q], n * sizeof(double), hipMemcpyHostToDevice, H->s));
                *in_dst[q] = P[q];
            }
        }
        if (dev) {
            A.iat = f->nu_ano_IAT;
        } else {
            A.iat = H->trp.p + NIN * n;
            CATHODE_HIP(hipMemcpyAsync(A.iat, f->nu_ano_IAT, n * sizeof(double), hipMemcpyHostToDevice, H->s));
        }
        for (int q = 0; q < NOUT; q++) {
            if (!out_dst[q]) A.out[q] = nullptr;
            else A.out[q] = dev ? out_dst[q] : H->trp.p + (NIN + 1 + q) * n;
        }
        // Mache is written on the plasma cells only: the others keep the caller's values
        if (!dev && f->Mache)
            CATHODE_HIP(hipMemcpyAsync(A.out[TO_MACH], f->Mache, n * sizeof(double), hipMemcpyHostToDevice, H->s));
        const int ne = H->rate_n[0];
        for (int q = 0; q < TB_COUNT; q++) A.tab[q] = H->rate_tab.p + (size_t)q * ne;
        A.nui_tab = H->rate_tab.p + (size_t)TB_COUNT * ne;
        A.n_e = ne;
        A.n_i = H->rate_n[1];
        A.dTe = H->rate_d[0];
        A.dEi = H->rate_d[1];
        A.masse = p->masse;
        A.massi = p->massi;
        A.Ti = p->Ti;
        A.nu_eff = p->nu_eff;
        A.flag_Bohm = (double)p->flag_Bohm;
        A.flag_Sagdeev = (double)p->flag_Sagdeev;
        H->tr_scal.alloc(1);
        A.rhoe_max = H->tr_scal.p;
        NDD_GEO_LAUNCH(*H, k_rhoe_max_part, ggrid(V.ni, V.nj), gblk(), V, A.rhoe, H->part.p);
        hipLaunchKernelGGL(k_rhoe_max_finish, dim3(1), dim3(256), 0, H->s, (const double*)H->part.p, H->nb,
                           H->tr_scal.p);
        CATHODE_LAUNCH_CHECK();
        NDD_GEO_LAUNCH(*H, k_transport, ggrid(V.ni + 2, V.nj + 2), gblk(), V, A);
        CATHODE_LAUNCH_CHECK();
        if (!dev) {
            CATHODE_HIP(hipMemcpyAsync(f->nu_ano_IAT, A.iat, n * sizeof(double), hipMemcpyDeviceToHost, H->s));
            for (int q = 0; q < NOUT; q++)
                if (out_dst[q])
                    CATHODE_HIP(hipMemcpyAsync(out_dst[q], A.out[q], n * sizeof(double), hipMemcpyDeviceToHost, H->s));
        }
        CATHODE_HIP(hipStreamSynchronize(H->s));
        return CATHODE_OK;
    CATHODE_CATCH_END
}

extern "C" int cathode_ndd_transport_update(void* handle, const cathode_transport_params* p,
                                            cathode_transport_fields* f)
{
    return transport_entry(handle, p, f, false);
}

extern "C" int cathode_ndd_transport_update_dev(void* handle, const cathode_transport_params* p,
                                                cathode_transport_fields* f)
{
    return transport_entry(handle, p, f, true);
}
