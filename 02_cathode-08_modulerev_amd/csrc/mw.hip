// mw.hip -- microwave power profile and port normalisation on the device:
// EmfieldModule::update_energy_profile (include/emfieldModule.hpp:7630-8105)
// with the Poynting fluxes of mWLoss_at_boundary_xL / rR (:8488-8668).
//
// J = sigma E with the magnetised cold-plasma conductivity on x faces, r faces
// and cells (each block's box in turn, as the reference's block loops), the
// absorbed power per component and its volume sums (fixed-order partials),
// the cell absorbed power; then the reference-plane current, voltage and
// Poynting fluxes (one workgroup), the port algebra on the host (a few complex
// scalars), and one elementwise launch that rescales Pabs, E, J and the
// excitation currents.  Face and cell conventions: x face i lies between cells
// i-1 and i, r face j between j-1 and j; r_j = rmin + (j - 1/2) dr.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "cplx.hpp"
#include "helm.hpp"

namespace cathode {

constexpr double MW_E0 = 1.60217662e-19;   // include/constants.hpp:6
constexpr double MW_MU0 = 1.2566370614e-6; // :8
constexpr int MX = 64, MY = 4;

struct MwView {
    int ni, nj, ld, n_bl;
    double dx, dr, omega, masse;
    int box[3][4 * CATHODE_MW_MAX_BLOCKS];  // cells, x faces, r faces
    const double* r;
};

struct Cond {
    cxd hx, hr, coef;
};

// conductivity factors of one point (emfieldModule.hpp:7648-7660)
__device__ inline Cond conductivity(double n, double nu, double Bx, double Br, double omega, double masse)
{
#pragma clang fp contract(off)
    const cxd EPS(1e-100, 1e-100);
    const cxd nus(nu, omega);
    const cxd den = cxd(masse) * nus + EPS;
    Cond c;
    c.hx = cxd(MW_E0 * Bx) / den;
    c.hr = cxd(MW_E0 * Br) / den;
    const cxd h2 = c.hx * c.hx + c.hr * c.hr;
    c.coef = cxd(n * MW_E0 * MW_E0) / den / (cxd(1.0) + h2 + EPS);
    return c;
}

__device__ inline double mw_block_sum(double v, double* lds)
{
    return block_sum<MX * MY>(v, lds);
}

// component k: 0 x faces, 1 r faces, 2 cells; blockIdx.z = block (box) index;
// partial volume sum of each workgroup into part[(z * gridDim.y + y) * gridDim.x + x]
template <int K>
__global__ void __launch_bounds__(MX* MY)
    k_mw_component(MwView V, const double* __restrict__ rhoe, const double* __restrict__ nu,
                   const double* __restrict__ Bx, const double* __restrict__ Br, const cxd* __restrict__ Ex,
                   const cxd* __restrict__ Er, const cxd* __restrict__ Ep, cxd* __restrict__ J,
                   double* __restrict__ Pab, double* __restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double lds[MX * MY / 64];
    const int* bx = V.box[K == 0 ? 1 : (K == 1 ? 2 : 0)] + 4 * blockIdx.z;
    const int i = bx[0] + (int)(blockIdx.y * MY + threadIdx.y);
    const int j = bx[2] + (int)(blockIdx.x * MX + threadIdx.x);
    const int ld = V.ld;
    const double* r = V.r;
    double pv = 0.0;
    if (i <= bx[1] && j <= bx[3]) {
        const size_t k = (size_t)i * ld + j;
        const double vol = V.dx * V.dr * 2.0 * M_PI;
        if constexpr (K == 0) {
            const double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
            const Cond c = conductivity((rhoe[k] + rhoe[k - ld]) / 2.0, (nu[k] + nu[k - ld]) / 2.0,
                                        (Bx[k] + Bx[k - ld]) / 2.0, (Br[k] + Br[k - ld]) / 2.0, V.omega, V.masse);
            const cxd sxr = c.coef * (c.hx * c.hr), sxp = c.coef * c.hr, sxx = c.coef * (cxd(1.0) + c.hx * c.hx);
            const cxd ex = Ex[k];
            const cxd er = (cxd(rL) * (Er[k] + Er[k - ld]) + cxd(rR) * (Er[k + 1] + Er[k - ld + 1])) / cxd(4.0 * r[j]);
            const cxd ep = (Ep[k] + Ep[k - ld]) / cxd(2.0);
            const cxd jj = sxr * er + sxp * ep + sxx * ex;
            J[k] = jj;
            const double pa = 0.5 * (jj * conj(ex)).re;
            Pab[k] = pa;
            pv = pa * (vol * r[j]);
        } else if constexpr (K == 1) {
            const double rt = (r[j - 1] + r[j]) / 2.0;
            const Cond c = conductivity((rhoe[k] + rhoe[k - 1]) / 2.0, (nu[k] + nu[k - 1]) / 2.0,
                                        (r[j] * Bx[k] + r[j - 1] * Bx[k - 1]) / (2.0 * rt),
                                        (r[j] * Br[k] + r[j - 1] * Br[k - 1]) / (2.0 * rt), V.omega, V.masse);
            const cxd srr = c.coef * (cxd(1.0) + c.hr * c.hr), srp = c.coef * (-c.hx), srx = c.coef * (c.hx * c.hr);
            const cxd ex = (cxd(r[j]) * (Ex[k] + Ex[k + ld]) + cxd(r[j - 1]) * (Ex[k - 1] + Ex[k + ld - 1])) /
                           cxd(4.0 * rt);
            const cxd er = Er[k];
            const cxd ep = (cxd(r[j]) * Ep[k] + cxd(r[j - 1]) * Ep[k - 1]) / cxd(2.0 * rt);
            const cxd jj = srr * er + srp * ep + srx * ex;
            J[k] = jj;
            const double pa = 0.5 * (jj * conj(er)).re;
            Pab[k] = pa;
            pv = pa * (vol * rt);
        } else {
            const double qL = (r[j - 1] + r[j]) / 2.0 / (r[j] + 1e-100);
            const double qR = (r[j + 1] + r[j]) / 2.0 / (r[j] + 1e-100);
            const Cond c = conductivity(rhoe[k], nu[k], Bx[k], Br[k], V.omega, V.masse);
            const cxd spr = c.coef * c.hx, spp = c.coef * cxd(1.0), spx = c.coef * (-c.hr);
            const cxd ex = (Ex[k] + Ex[k + ld]) / cxd(2.0);
            const cxd er = (cxd(qR) * Er[k + 1] + cxd(qL) * Er[k]) / cxd(2.0);
            const cxd ep = Ep[k];
            const cxd jj = spr * er + spp * ep + spx * ex;
            J[k] = jj;
            const double pa = 0.5 * (jj * conj(ep)).re;
            Pab[k] = pa;
            pv = pa * (vol * r[j]);
        }
    }
    const double s = mw_block_sum(pv, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// cell absorbed power: the face averages plus the phi part (:7750-7790)
__global__ void k_mw_cell_pabs(MwView V, const double* __restrict__ Px, const double* __restrict__ Pr,
                               const double* __restrict__ Pp, double* __restrict__ P)
{
#pragma clang fp contract(off)
    const int* bx = V.box[0] + 4 * blockIdx.z;
    const int i = bx[0] + (int)(blockIdx.y * MY + threadIdx.y);
    const int j = bx[2] + (int)(blockIdx.x * MX + threadIdx.x);
    if (i > bx[1] || j > bx[3]) return;
    const size_t k = (size_t)i * V.ld + j;
    P[k] = (Px[k] + Px[k + V.ld]) / 2.0 + (Pr[k] + Pr[k + 1]) / 2.0 + Pp[k];
}

struct MwRef {
    int nparts[3];    // partial counts of the three components
    int i_ref, jf_lo, jf_hi;        // reference plane x face, the feed's r faces
    int j_out, i_out_lo, i_out_hi;  // loss boundary
    int j_in_lo, j_in_hi;           // Poynting input cells
};

// out: [0..2] Px Pr Pp, [3] P_out, [4] P_in, [5..6] I_ref sum (re, im), [7..8]
// sum of (E1r(i) + E1r(i-1)) / 2 * r over the feed's faces, [9] their count.
// One workgroup of 256 threads; every sum in a fixed order.
__global__ void __launch_bounds__(256)
    k_mw_reference(MwView V, MwRef R, const double* __restrict__ part, const cxd* __restrict__ Ex,
                   const cxd* __restrict__ Er, const cxd* __restrict__ Ep, double* __restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ double lds[8];
    const int t = threadIdx.x, ld = V.ld;
    const double* r = V.r;
    const double* pp = part;
    for (int c = 0; c < 3; c++) {
        double s = 0.0;
        for (int q = t; q < R.nparts[c]; q += 256) s += pp[q];
        s = block_sum<256>(s, lds);
        if (t == 0) out[c] = s;
        pp += R.nparts[c];
    }
    const cxd iwm(0.0, 1.0 / (V.omega * MW_MU0));
    // loss through the r face above cells (i, j_out), i = i_out_lo .. i_out_hi (mWLoss_at_boundary_rR)
    double pl = 0.0;
    for (int i = R.i_out_lo + t; i <= R.i_out_hi; i += 256) {
        const int j = R.j_out;
        const double rt = (r[j] + r[j + 1]) / 2.0;
        auto at = [&](const cxd* a, int ii, int jj) { return a[(size_t)ii * ld + jj]; };
        const cxd dErdx_L = (at(Er, i, j + 1) - at(Er, i - 1, j + 1)) / cxd(V.dx);
        const cxd dExdr_L = (at(Ex, i, j + 1) - at(Ex, i, j)) / cxd(V.dr);
        const cxd dErdx_R = (at(Er, i + 1, j + 1) - at(Er, i, j + 1)) / cxd(V.dx);
        const cxd dExdr_R = (at(Ex, i + 1, j + 1) - at(Ex, i + 1, j)) / cxd(V.dr);
        const cxd Hp = (iwm * (dErdx_R - dExdr_R) + iwm * (dErdx_L - dExdr_L)) / cxd(2.0);
        const cxd dEpdr = (cxd(r[j + 1]) * at(Ep, i, j + 1) - cxd(r[j]) * at(Ep, i, j)) / cxd(rt * V.dr);
        const cxd Hx = iwm * dEpdr;
        const cxd Epf = (cxd(r[j + 1]) * at(Ep, i, j + 1) + cxd(r[j]) * at(Ep, i, j)) / cxd(2.0 * rt);
        const cxd ExL = (cxd(r[j + 1]) * at(Ex, i, j + 1) + cxd(r[j]) * at(Ex, i, j)) / cxd(2.0 * rt);
        const cxd ExR = (cxd(r[j + 1]) * at(Ex, i + 1, j + 1) + cxd(r[j]) * at(Ex, i + 1, j)) / cxd(2.0 * rt);
        const cxd Exf = (ExR + ExL) / cxd(2.0);
        const cxd ExH = Epf * conj(Hx) - Exf * conj(Hp);
        pl += 2.0 * M_PI * rt * (0.5 * ExH.re) * V.dx;
    }
    pl = block_sum<256>(pl, lds);
    // input through the x face i_ref over cells j_in_lo .. j_in_hi (mWLoss_at_boundary_xL)
    double pin = 0.0;
    for (int j = R.j_in_lo + t; j <= R.j_in_hi; j += 256) {
        const int i = R.i_ref;
        const double rL = (r[j] + r[j - 1]) / 2.0, rR = (r[j] + r[j + 1]) / 2.0;
        auto at = [&](const cxd* a, int ii, int jj) { return a[(size_t)ii * ld + jj]; };
        const cxd dErdx_L = (at(Er, i, j) - at(Er, i - 1, j)) / cxd(V.dx);
        const cxd dExdr_L = (at(Ex, i, j) - at(Ex, i, j - 1)) / cxd(V.dr);
        const cxd dErdx_R = (at(Er, i, j + 1) - at(Er, i - 1, j + 1)) / cxd(V.dx);
        const cxd dExdr_R = (at(Ex, i, j + 1) - at(Ex, i, j)) / cxd(V.dr);
        const cxd Hp = (cxd(rR) * (iwm * (dErdx_R - dExdr_R)) + cxd(rL) * (iwm * (dErdx_L - dExdr_L))) /
                       cxd(2.0 * r[j]);
        const cxd Epf = (at(Ep, i, j) + at(Ep, i - 1, j)) / cxd(2.0 * rR);
        const cxd ErL = (cxd(rR) * at(Er, i, j + 1) + cxd(rL) * at(Er, i, j)) / cxd(2.0 * r[j]);
        const cxd ErR = (cxd(rR) * at(Er, i - 1, j + 1) + cxd(rL) * at(Er, i - 1, j)) / cxd(2.0 * r[j]);
        const cxd Erf = (ErR + ErL) / cxd(2.0);
        const cxd ExH = Erf * conj(Hp) - Epf * conj(Hp);
        pin += 2.0 * M_PI * r[j] * (0.5 * ExH.re) * V.dr;
    }
    pin = block_sum<256>(pin, lds);
    // reference-plane current and voltage sums over the feed's r faces (:7874-7920)
    double ire = 0.0, iim = 0.0, cre = 0.0, cim = 0.0;
    for (int j = R.jf_lo + t; j <= R.jf_hi; j += 256) {
        const int i = R.i_ref;
        const double rt = (r[j] + r[j - 1]) / 2.0;
        const cxd e1 = Er[(size_t)i * ld + j], e0 = Er[(size_t)(i - 1) * ld + j];
        const cxd H = (e1 - e0) / cxd(V.dx) / cxd(V.omega * MW_MU0) * cxd(0.0, 1.0);
        const cxd Ii = cxd(2.0 * M_PI * rt) * H;
        const cxd c = (e1 + e0) / cxd(2.0) * cxd(rt);
        ire += Ii.re;
        iim += Ii.im;
        cre += c.re;
        cim += c.im;
    }
    ire = block_sum<256>(ire, lds);
    iim = block_sum<256>(iim, lds);
    cre = block_sum<256>(cre, lds);
    cim = block_sum<256>(cim, lds);
    if (t == 0) {
        out[3] = pl;
        out[4] = pin;
        out[5] = ire;
        out[6] = iim;
        out[7] = cre;
        out[8] = cim;
        out[9] = (double)std::max(0, R.jf_hi - R.jf_lo + 1);
    }
}

// the normalisation: Pabs *= pscale everywhere; with fscale > 0 the fields,
// currents and excitation currents *= fscale
__global__ void k_mw_scale(size_t n, double pscale, double fscale, double* __restrict__ Pabs, cxd* E1, cxd* E2,
                           cxd* E3, cxd* J1, cxd* J2, cxd* J3, cxd* X1, cxd* X2, cxd* X3)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    Pabs[k] = Pabs[k] * pscale;
    if (fscale > 0.0) {
        const cxd f(fscale);
        E1[k] = E1[k] * f;
        E2[k] = E2[k] * f;
        E3[k] = E3[k] * f;
        J1[k] = J1[k] * f;
        J2[k] = J2[k] * f;
        J3[k] = J3[k] * f;
        if (X1) X1[k] = X1[k] * f;
        if (X2) X2[k] = X2[k] * f;
        if (X3) X3[k] = X3[k] * f;
    }
}

// device buffers of the last shape (calls are serialised by the API mutex)
struct MwBuffers {
    int ni = -1, nj = -1;
    DevBuf<double> real;   // rhoe, nu, Bx, Br, Pabsx, Pabsr, Pabsp, Pabs, radii
    DevBuf<cxd> cplx;      // E1x, E1r, E1p, J1x, J1r, J1p, J1x_exc, J1r_exc, J1p_exc
    DevBuf<double> part, out;
};
static MwBuffers g_mw;

void mw_release()
{
    g_mw.real.release();
    g_mw.cplx.release();
    g_mw.part.release();
    g_mw.out.release();
    g_mw.ni = g_mw.nj = -1;
}

static bool box_ok(const int* b, int ilo, int ihi, int jlo, int jhi)
{
    return b[0] >= ilo && b[1] <= ihi && b[2] >= jlo && b[3] <= jhi;
}

}  // namespace cathode

using namespace cathode;

extern "C" int cathode_mw_energy_profile(const cathode_mw_params* p, cathode_mw_fields* f, cathode_mw_report* rep)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    if (!p || !f || !rep || p->ni < 2 || p->nj < 2 || p->n_bl < 1 || p->n_bl > CATHODE_MW_MAX_BLOCKS ||
        !(p->dx > 0) || !(p->dr > 0) || !(p->omega > 0) || !(p->masse > 0) || !f->rhoe || !f->nu_m || !f->Bx ||
        !f->Br || !f->E1x || !f->E1r || !f->E1p || !f->J1x || !f->J1r || !f->J1p || !f->Pabsx || !f->Pabsr ||
        !f->Pabsp || !f->Pabs) {
        set_error("mw_energy_profile: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    const int ni = p->ni, nj = p->nj, ld = nj + 2;
    // every box must leave the one-cell stencil inside the extended grid
    for (int b = 0; b < p->n_bl; b++)
        if (!box_ok(p->cell_box + 4 * b, 1, ni, 1, nj) || !box_ok(p->xface_box + 4 * b, 1, ni + 1, 1, nj) ||
            !box_ok(p->rface_box + 4 * b, 1, ni, 1, nj + 1)) {
            set_error("mw_energy_profile: block boxes must lie inside cells 1..ni x 1..nj (x faces 1..ni+1, "
                      "r faces 1..nj+1)");
            return CATHODE_ERR_ARG;
        }
    if (p->flag_mwRef == 1) {
        const int* fb = p->rface_box + 4 * (p->n_bl - 1);
        if (p->i_ref < 1 || p->i_ref > ni + 1 || p->j_out < 1 || p->j_out > nj || p->i_out_lo < 1 ||
            p->i_out_hi > ni || p->j_in_lo < 1 || p->j_in_hi > nj || fb[2] > fb[3] || !(p->r2 > p->r1) ||
            !(p->r1 > 0) || !(p->Z0 > 0)) {
            set_error("mw_energy_profile: reference plane, loss boundary or feed radii out of range");
            return CATHODE_ERR_ARG;
        }
    }
    hipStream_t s = ctx().stream;
    const size_t n = (size_t)(ni + 2) * ld;
    MwBuffers& B = g_mw;
    if (B.ni != ni || B.nj != nj) {
        B.real.alloc(8 * n + ld);
        B.cplx.alloc(9 * n);
        CATHODE_HIP(hipMemsetAsync(B.real.p, 0, (8 * n + ld) * sizeof(double), s));
        CATHODE_HIP(hipMemsetAsync(B.cplx.p, 0, 9 * n * sizeof(cxd), s));
        B.ni = ni;
        B.nj = nj;
    }
    double* R[8];
    for (int q = 0; q < 8; q++) R[q] = B.real.p + q * n;
    double* radii = B.real.p + 8 * n;
    cxd* C[9];
    for (int q = 0; q < 9; q++) C[q] = B.cplx.p + q * n;
    MwView V{};
    V.ni = ni; V.nj = nj; V.ld = ld; V.n_bl = p->n_bl;
    V.dx = p->dx; V.dr = p->dr; V.omega = p->omega; V.masse = p->masse;
    std::memcpy(V.box[0], p->cell_box, sizeof(p->cell_box));
    std::memcpy(V.box[1], p->xface_box, sizeof(p->xface_box));
    std::memcpy(V.box[2], p->rface_box, sizeof(p->rface_box));
    V.r = radii;
    std::vector<double> r(ld);
    for (int j = 0; j < ld; j++) r[j] = p->rmin + (j - 0.5) * p->dr;  // initialModule.hpp:86-88
    CATHODE_HIP(hipMemcpyAsync(radii, r.data(), ld * sizeof(double), hipMemcpyHostToDevice, s));
    // every array is uploaded: the outputs keep their values outside the
    // blocks' boxes (and are rescaled there too), as the reference's members do
    const double* rin[8] = {f->rhoe, f->nu_m, f->Bx, f->Br, f->Pabsx, f->Pabsr, f->Pabsp, f->Pabs};
    for (int q = 0; q < 8; q++)
        CATHODE_HIP(hipMemcpyAsync(R[q], rin[q], n * sizeof(double), hipMemcpyHostToDevice, s));
    double* cio[9] = {f->E1x, f->E1r, f->E1p, f->J1x, f->J1r, f->J1p, f->J1x_exc, f->J1r_exc, f->J1p_exc};
    for (int q = 0; q < 9; q++)
        if (cio[q]) CATHODE_HIP(hipMemcpyAsync(C[q], cio[q], n * sizeof(cxd), hipMemcpyHostToDevice, s));
    // grids: the largest box of each kind; a workgroup's partial slot per (box, block)
    dim3 g3[3];
    int nparts[3];
    for (int c = 0; c < 3; c++) {
        const int* bb = c == 0 ? p->xface_box : (c == 1 ? p->rface_box : p->cell_box);
        int mi = 1, mj = 1;
        for (int b = 0; b < p->n_bl; b++) {
            mi = std::max(mi, bb[4 * b + 1] - bb[4 * b] + 1);
            mj = std::max(mj, bb[4 * b + 3] - bb[4 * b + 2] + 1);
        }
        g3[c] = dim3(cdiv(mj, MX), cdiv(mi, MY), p->n_bl);
        nparts[c] = (int)(g3[c].x * g3[c].y * g3[c].z);
    }
    B.part.alloc((size_t)nparts[0] + nparts[1] + nparts[2] + 64);
    B.out.alloc(16);
    double* pp = B.part.p;
    hipLaunchKernelGGL(k_mw_component<0>, g3[0], dim3(MX, MY), 0, s, V, R[0], R[1], R[2], R[3], C[0], C[1], C[2],
                       C[3], R[4], pp);
    hipLaunchKernelGGL(k_mw_component<1>, g3[1], dim3(MX, MY), 0, s, V, R[0], R[1], R[2], R[3], C[0], C[1], C[2],
                       C[4], R[5], pp + nparts[0]);
    hipLaunchKernelGGL(k_mw_component<2>, g3[2], dim3(MX, MY), 0, s, V, R[0], R[1], R[2], R[3], C[0], C[1], C[2],
                       C[5], R[6], pp + nparts[0] + nparts[1]);
    // the cell absorbed power reads every component's faces: after all three
    hipLaunchKernelGGL(k_mw_cell_pabs, g3[2], dim3(MX, MY), 0, s, V, (const double*)R[4], (const double*)R[5],
                       (const double*)R[6], R[7]);
    MwRef ref{};
    ref.nparts[0] = nparts[0];
    ref.nparts[1] = nparts[1];
    ref.nparts[2] = nparts[2];
    if (p->flag_mwRef == 1) {
        const int* fb = p->rface_box + 4 * (p->n_bl - 1);
        ref.i_ref = p->i_ref;
        ref.jf_lo = fb[2];
        ref.jf_hi = fb[3];
        ref.j_out = p->j_out;
        ref.i_out_lo = p->i_out_lo;
        ref.i_out_hi = p->i_out_hi;
        ref.j_in_lo = p->j_in_lo;
        ref.j_in_hi = p->j_in_hi;
    } else {
        // no plane or boundary: the loops run empty
        ref.i_ref = 1;
        ref.jf_lo = 1;
        ref.jf_hi = 0;
        ref.j_out = 1;
        ref.i_out_lo = 1;
        ref.i_out_hi = 0;
        ref.j_in_lo = 1;
        ref.j_in_hi = 0;
    }
    hipLaunchKernelGGL(k_mw_reference, dim3(1), dim3(256), 0, s, V, ref, (const double*)pp, (const cxd*)C[0],
                       (const cxd*)C[1], (const cxd*)C[2], B.out.p);
    CATHODE_LAUNCH_CHECK();
    double h[16];
    CATHODE_HIP(hipMemcpyAsync(h, B.out.p, 10 * sizeof(double), hipMemcpyDeviceToHost, s));
    CATHODE_HIP(hipStreamSynchronize(s));
    // the port algebra of the reference (:7920-8060), a few complex scalars
    std::memset(rep, 0, sizeof(*rep));
    typedef std::complex<double> C128;
    const double Px = h[0], Pr = h[1], Pp = h[2];
    double power_sum = Px + Pr + Pp;
    rep->Px = Px;
    rep->Pr = Pr;
    rep->Pp = Pp;
    double pscale = 1.0, fscale = 0.0;
    if (p->flag_mwRef == 1) {
        const double P_out = h[3], P_in = h[4];
        const double cnt = h[9];
        const C128 Iref = C128(h[5], h[6]) / cnt;
        const C128 Vref = C128(h[7], h[8]) / cnt * std::log(p->r2 / p->r1);
        const C128 Zref = Vref / (Iref + 1e-100);
        const double Z0 = p->Z0;
        const C128 a2 = (Vref - Z0 * Iref) / (2.0 * std::sqrt(2.0 * Z0));
        const C128 b2 = (Vref + Z0 * Iref) / (2.0 * std::sqrt(2.0 * Z0));
        const double P_fwd = std::norm(b2), P_ref = std::norm(a2);
        const double P_err = P_fwd - P_ref - power_sum - P_out;
        const C128 S11(p->S11[0], p->S11[1]), S12(p->S12[0], p->S12[1]), S21(p->S21[0], p->S21[1]),
            S22(p->S22[0], p->S22[1]);
        const C128 a1 = (b2 - S22 * a2) / S21;
        const C128 b1 = S11 * a1 + S12 * a2;
        const double P_fwd_all = std::norm(a1), P_ref_all = std::norm(b1);
        const double ratio = p->Pmw / P_fwd_all;
        const double ratio_eff = ratio * (power_sum + P_err) / power_sum;
        pscale = ratio_eff;
        fscale = std::sqrt(ratio_eff);
        rep->J1r_exc = p->J1r_exc * std::sqrt(ratio);
        power_sum = power_sum + P_err;
        rep->P_fwd = P_fwd * ratio;
        rep->P_ref = P_ref * ratio;
        rep->P_fwd_all = P_fwd_all * ratio;
        rep->P_ref_all = P_ref_all * ratio;
        rep->power_sum = power_sum * ratio;
        rep->P_err = P_err * ratio;
        rep->P_out = P_out * ratio;
        rep->P_in = P_in;
        rep->ratio = ratio;
        rep->ratio_eff = ratio_eff;
        const C128 Gp = a2 / b2, Zp = Z0 * (1.0 + Gp) / (1.0 - Gp);
        const C128 Ga = b1 / a1, Za = Z0 * (1.0 + Ga) / (1.0 - Ga);
        const C128 vals[7] = {Vref, Iref, Zref, Gp, Zp, Ga, Za};
        double* dst[7] = {rep->V_ref, rep->I_ref, rep->Z_ref, rep->Gamma_p, rep->Zp, rep->Gamma_all, rep->Zall};
        for (int q = 0; q < 7; q++) {
            dst[q][0] = vals[q].real();
            dst[q][1] = vals[q].imag();
        }
    } else {
        const double ratio = p->Pmw / power_sum;
        pscale = ratio;
        rep->J1r_exc = p->J1r_exc * std::sqrt(ratio);
        rep->ratio = ratio;
        rep->ratio_eff = ratio;
        rep->power_sum = power_sum;
    }
    hipLaunchKernelGGL(k_mw_scale, dim3(cdiv((long)n, 256)), dim3(256), 0, s, n, pscale, fscale, R[7], C[0], C[1],
                       C[2], C[3], C[4], C[5], f->J1x_exc ? C[6] : nullptr, f->J1r_exc ? C[7] : nullptr,
                       f->J1p_exc ? C[8] : nullptr);
    CATHODE_LAUNCH_CHECK();
    double* rout[4] = {f->Pabsx, f->Pabsr, f->Pabsp, f->Pabs};
    for (int q = 0; q < 4; q++)
        CATHODE_HIP(hipMemcpyAsync(rout[q], R[4 + q], n * sizeof(double), hipMemcpyDeviceToHost, s));
    for (int q = 0; q < 9; q++)
        if (q < 6 || cio[q]) CATHODE_HIP(hipMemcpyAsync(cio[q], C[q], n * sizeof(cxd), hipMemcpyDeviceToHost, s));
    CATHODE_HIP(hipStreamSynchronize(s));
    return CATHODE_OK;
    CATHODE_GUARD_END
}
