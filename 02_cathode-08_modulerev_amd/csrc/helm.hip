// helm.hip -- complex 5-point solves: the E_theta microwave system.
//
// EmfieldModule::solve_Microwave assembles the (E_x, E_r, E_theta) system and
// factorizes it with Eigen SparseLU (include/emfieldModule.hpp:2828-2998).
// With B = 0 the E_theta rows decouple (:2414-2454): a complex r-weighted
// Helmholtz operator aP = 2dx^2 + 2dr^2 + dx^2dr^2/r^2 - eps0 epsr mu0 w^2
// dx^2dr^2 + i w mu0 sigma dx^2dr^2, aE = aW = dr^2, aN = qR dx^2,
// aS = qL dx^2, b = -i w mu0 dx^2dr^2 J.  This file solves it with the same
// multigrid-preconditioned BiCGStab as the real path, on complex values:
//   complex128 (mode FP64), or complex64 V-cycles with complex128 residuals
//   and BiCGStab vectors (mode MIXED).
// Coarse operators: as mg.hpp, with the row excess s = aP - sum(a_nb) kept
// complex (no clipping) and the "mass" part taken from the neighbour of
// smallest |s| (the volumetric 1/r^2 - k^2 + i w mu0 sigma term; PEC folds are
// the boundary part).  Inner products are conjugated: <a,b> = sum conj(a) b.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "cplx.hpp"
#include "helm.hpp"
#include "mg.hpp"

namespace cathode {
__device__ inline double abs2v(double v) { return v * v; }
__device__ inline float abs2v(float v) { return v * v; }
}  // namespace cathode
#include "vcycle.hpp"

namespace cathode {

constexpr int HX = 64, HY = 4;
constexpr int HT_CFG = 2;  // complex tiles: 8 x 32 cells, 256 threads (register budget)
constexpr int HL_MAX = 24;

static inline dim3 hgrid(int ni, int nj) { return dim3(cdiv(nj, HX), cdiv(ni, HY)); }
#define HCELL                                        \
    const int j = 1 + blockIdx.x * HX + threadIdx.x; \
    const int i = 1 + blockIdx.y * HY + threadIdx.y;

enum { HS_BB = 0, HS_SCALE = 1, HS_S32 = 2, HS_RR = 3, HS_RT = 4, HS_RHO = 5, HS_RHO1 = 7, HS_AL = 9, HS_OM = 11, HS_COUNT = 16 };
// complex scalars occupy two slots (re, im)
__device__ inline cxd ldc(const double* s, int k) { return cxd(s[k], s[k + 1]); }
__device__ inline void stc(double* s, int k, cxd v)
{
    s[k] = v.re;
    s[k + 1] = v.im;
}

// ------------------------------------------------------------------ setup kernels
__global__ void hk_prepare(int ni, int nj, int ld, const unsigned char* __restrict__ act, cxd* aP, cxd* aE, cxd* aW,
                           cxd* aN, cxd* aS, const cxd* __restrict__ b, cxd* __restrict__ sx,
                           double* __restrict__ part, int nb)
{
    __shared__ double lds[8];
    HCELL;
    double amax = 0.0, bb = 0.0;
    if (i <= ni && j <= nj) {
        const size_t k = (size_t)i * ld + j;
        if (!act[k]) {
            aP[k] = cxd(1.0);
            aE[k] = aW[k] = aN[k] = aS[k] = cxd(0.0);
            sx[k] = cxd(0.0);
        } else {
            const cxd e = act[k + ld] ? aE[k] : cxd(0.0), w = act[k - ld] ? aW[k] : cxd(0.0);
            const cxd n = act[k + 1] ? aN[k] : cxd(0.0), s = act[k - 1] ? aS[k] : cxd(0.0);
            aE[k] = e;
            aW[k] = w;
            aN[k] = n;
            aS[k] = s;
            sx[k] = aP[k] - (e + w + n + s);
            amax = sqrt(abs2v(aP[k]));
            bb = abs2v(b[k]);
        }
    }
    const double t1 = block_max<HX * HY>(amax, lds);
    const double t2 = block_sum<HX * HY>(bb, lds);
    if (lin_tid() == 0) {
        const int bid = blockIdx.y * gridDim.x + blockIdx.x;
        part[bid] = t1;
        part[nb + bid] = t2;
    }
}

__global__ void hk_setup_finish(const double* __restrict__ part, int nb, double* __restrict__ sc, double nact)
{
    __shared__ double lds[8];
    double m = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        m = fmax(m, part[k]);
        b += part[nb + k];
    }
    m = block_max<256>(m, lds);
    b = block_sum<256>(b, lds);
    if (threadIdx.x == 0) {
        sc[HS_BB] = b;
        sc[HS_SCALE] = b > 0.0 ? sqrt(nact / b) : 1.0;
        int e = 0;
        (void)frexp(m > 0.0 ? m : 1.0, &e);
        sc[HS_S32] = ldexp(1.0, -e);
    }
}

__global__ void hk_weights(int ni, int nj, int ld, const cxd* __restrict__ aN, const cxd* __restrict__ aS,
                           double* __restrict__ w)
{
    extern __shared__ double ratio[];
    for (int j = 1 + threadIdx.x; j < nj; j += blockDim.x) {
        double sN = 0.0, sS = 0.0;
        for (int i = 1; i <= ni; i++) {
            sN += aN[(size_t)i * ld + j].re;
            sS += aS[(size_t)i * ld + j + 1].re;
        }
        const double r = (sN > 0.0 && sS > 0.0) ? sN / sS : 1.0;
        ratio[j] = r < 1e-2 ? 1e-2 : (r > 1e2 ? 1e2 : r);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 1.0;
        w[0] = w[1] = 1.0;
        for (int j = 1; j < nj; j++) {
            v *= ratio[j];
            w[j + 1] = v;
        }
        w[nj + 1] = v;
    }
}

// mass part: w_j times the excess of smallest modulus in the 5-point
// neighbourhood; FP32 copy of level 0 in mixed mode
__global__ void hk_mass(int ni, int nj, int ld, const unsigned char* __restrict__ act, const cxd* __restrict__ sx,
                        const double* __restrict__ w, cxd* __restrict__ m, const cxd* __restrict__ aP,
                        const cxd* __restrict__ aE, const cxd* __restrict__ aW, const cxd* __restrict__ aN,
                        const cxd* __restrict__ aS, cxf* fP, cxf* fE, cxf* fW, cxf* fN, cxf* fS,
                        const double* __restrict__ sc, int mixed)
{
    HCELL;
    if (i > ni || j > nj) return;
    const size_t k = (size_t)i * ld + j;
    if (mixed) {
        const double s32 = sc[HS_S32];
        const cxd q(s32, 0.0);
        fP[k] = cxf(aP[k] * q);
        fE[k] = cxf(aE[k] * q);
        fW[k] = cxf(aW[k] * q);
        fN[k] = cxf(aN[k] * q);
        fS[k] = cxf(aS[k] * q);
    }
    if (!act[k]) {
        m[k] = cxd(0.0);
        return;
    }
    cxd v = sx[k];
    const size_t nbk[4] = {k + (size_t)ld, k - (size_t)ld, k + 1, k - 1};
    for (int t = 0; t < 4; t++)
        if (act[nbk[t]] && abs2v(sx[nbk[t]]) < abs2v(v)) v = sx[nbk[t]];
    m[k] = cxd(w[j] * v.re, w[j] * v.im);
}

struct HCoarsen {
    int fni, fnj, fld, fi, fj;
    const cxd *faP, *faE, *faW, *faN, *faS, *fm;
    const unsigned char* fact;
    const double* w;
    int cni, cnj, cld;
    cxd *caP, *caE, *caW, *caN, *caS, *cb, *cx, *cm;
    unsigned char* cact;
    cxf *gP, *gE, *gW, *gN, *gS;
    double hx, hy, hb;
};

__device__ inline void hcoarsen_cell(const HCoarsen& a, int I, int J, const double* sc, int mixed)
{
    const int i0 = a.fi * (I - 1) + 1, i1 = min(a.fi * I, a.fni);
    const int j0 = a.fj * (J - 1) + 1, j1 = min(a.fj * J, a.fnj);
    cxd S(0.0), M(0.0), E(0.0), W(0.0), N(0.0), So(0.0);
    int any = 0;
    for (int i = i0; i <= i1; i++)
        for (int j = j0; j <= j1; j++) {
            const size_t k = (size_t)i * a.fld + j;
            if (!a.fact[k]) continue;
            any = 1;
            const cxd wj(a.w ? a.w[j] : 1.0, 0.0);
            const cxd e = a.faE[k], ww = a.faW[k], n = a.faN[k], s = a.faS[k];
            S += wj * (a.faP[k] - (e + ww + n + s));
            M += a.fm[k];
            if (i == i1) E += wj * e;
            if (i == i0) W += wj * ww;
            if (j == j1) N += wj * n;
            if (j == j0) So += wj * s;
        }
    const size_t kc = (size_t)I * a.cld + J;
    a.cact[kc] = (unsigned char)any;
    a.cb[kc] = cxd(0.0);
    a.cx[kc] = cxd(0.0);
    cxd P(1.0);
    if (!any) {
        E = W = N = So = cxd(0.0);
        a.cm[kc] = cxd(0.0);
    } else {
        E = E * cxd(a.hx);
        W = W * cxd(a.hx);
        N = N * cxd(a.hy);
        So = So * cxd(a.hy);
        a.cm[kc] = M;
        P = (M + cxd(a.hb) * (S - M)) + (E + W + N + So);
    }
    a.caP[kc] = P;
    a.caE[kc] = E;
    a.caW[kc] = W;
    a.caN[kc] = N;
    a.caS[kc] = So;
    if (mixed) {
        const cxd q(sc[HS_S32], 0.0);
        a.gP[kc] = cxf(P * q);
        a.gE[kc] = cxf(E * q);
        a.gW[kc] = cxf(W * q);
        a.gN[kc] = cxf(N * q);
        a.gS[kc] = cxf(So * q);
    }
}

__global__ void hk_coarsen(HCoarsen a, const double* __restrict__ sc, int mixed)
{
    const int J = 1 + blockIdx.x * HX + threadIdx.x;
    const int I = 1 + blockIdx.y * HY + threadIdx.y;
    if (I > a.cni || J > a.cnj) return;
    hcoarsen_cell(a, I, J, sc, mixed);
}

// ------------------------------------------------------------------ BiCGStab kernels
template <class T>
__device__ inline T hto_pc(cxd v, const double* sc, int mixed)
{
    if (!mixed) return T(v);
    const double f = sc[HS_S32] * sc[HS_SCALE];
    return T(cxd(v.re * f, v.im * f));
}
template <class T>
__device__ inline cxd hfrom_pc(T y, const double* sc, int mixed)
{
    const cxd v(y);
    if (!mixed) return v;
    const double f = 1.0 / sc[HS_SCALE];
    return cxd(v.re * f, v.im * f);
}

__device__ inline void block_store3(double a, double b, double c, double* part, int nb, double* lds)
{
    const double t1 = block_sum<HX * HY>(a, lds);
    const double t2 = block_sum<HX * HY>(b, lds);
    const double t3 = block_sum<HX * HY>(c, lds);
    if (lin_tid() == 0) {
        const int bid = blockIdx.y * gridDim.x + blockIdx.x;
        part[bid] = t1;
        part[nb + bid] = t2;
        part[2 * nb + bid] = t3;
    }
}

__device__ inline cxd hAx(const cxd* aP, const cxd* aE, const cxd* aW, const cxd* aN, const cxd* aS, const cxd* x,
                          size_t k, int ld)
{
    return aP[k] * x[k] - aE[k] * x[k + ld] - aW[k] * x[k - ld] - aN[k] * x[k + 1] - aS[k] * x[k - 1];
}

struct HOp {
    int ni, nj, ld;
    const cxd *aP, *aE, *aW, *aN, *aS;
    const unsigned char* act;
};

// r = b - A x, r0 = r, p = v = 0; partials |r|^2, <r0,r> (re, im)
__global__ void hk_restart(HOp A, const cxd* __restrict__ b, const cxd* __restrict__ x, cxd* r, cxd* r0, cxd* p,
                           cxd* v, double* part, int nb)
{
    __shared__ double lds[8];
    HCELL;
    double rr = 0.0;
    if (i <= A.ni && j <= A.nj) {
        const size_t k = (size_t)i * A.ld + j;
        cxd q(0.0);
        if (A.act[k]) {
            q = b[k] - hAx(A.aP, A.aE, A.aW, A.aN, A.aS, x, k, A.ld);
            rr = abs2v(q);
        }
        r[k] = q;
        r0[k] = q;
        p[k] = cxd(0.0);
        v[k] = cxd(0.0);
    }
    block_store3(rr, rr, 0.0, part, nb, lds);
}

template <class T>
__global__ void hk_p(HOp A, const cxd* __restrict__ r, const cxd* __restrict__ v, cxd* p, T* vin, T* vx,
                     const double* __restrict__ sc, int first, int mixed)
{
    HCELL;
    if (i > A.ni || j > A.nj) return;
    const size_t k = (size_t)i * A.ld + j;
    cxd pk(0.0);
    if (A.act[k]) {
        if (first) {
            pk = r[k];
        } else {
            cxd beta = (ldc(sc, HS_RHO1) / ldc(sc, HS_RHO)) * (ldc(sc, HS_AL) / ldc(sc, HS_OM));
            if (!(isfinite(beta.re) && isfinite(beta.im))) beta = cxd(0.0);  // breakdown: restart from r
            pk = r[k] + beta * (p[k] - ldc(sc, HS_OM) * v[k]);
        }
    }
    p[k] = pk;
    vin[k] = hto_pc<T>(pk, sc, mixed);
    vx[k] = T(0);
}

// y = M^-1 in (from the V-cycle output), out = A y; conj_out = 0: partial
// <dotvec, out>; 1: <out, dotvec> and |out|^2
template <class T>
__global__ void hk_apply(HOp A, const T* __restrict__ vx, cxd* y, cxd* out, const cxd* __restrict__ dotvec,
                         const double* __restrict__ sc, int mixed, int conj_out, double* part, int nb)
{
    __shared__ double lds[8];
    HCELL;
    double dre = 0.0, dim = 0.0, oo = 0.0;
    if (i <= A.ni && j <= A.nj) {
        const size_t k = (size_t)i * A.ld + j;
        cxd o(0.0), yk(0.0);
        if (A.act[k]) {
            const int ld = A.ld;
            yk = hfrom_pc<T>(vx[k], sc, mixed);
            o = A.aP[k] * yk - A.aE[k] * hfrom_pc<T>(vx[k + ld], sc, mixed) - A.aW[k] * hfrom_pc<T>(vx[k - ld], sc, mixed) -
                A.aN[k] * hfrom_pc<T>(vx[k + 1], sc, mixed) - A.aS[k] * hfrom_pc<T>(vx[k - 1], sc, mixed);
            const cxd d = conj_out ? conj(o) * dotvec[k] : conj(dotvec[k]) * o;
            dre = d.re;
            dim = d.im;
            oo = abs2v(o);
        }
        y[k] = yk;
        out[k] = o;
    }
    block_store3(dre, dim, oo, part, nb, lds);
}

template <class T>
__global__ void hk_s(HOp A, const cxd* __restrict__ r, const cxd* __restrict__ v, cxd* sv, T* vin, T* vx,
                     const double* __restrict__ sc, int mixed)
{
    HCELL;
    if (i > A.ni || j > A.nj) return;
    const size_t k = (size_t)i * A.ld + j;
    const cxd q = A.act[k] ? r[k] - ldc(sc, HS_AL) * v[k] : cxd(0.0);
    sv[k] = q;
    vin[k] = hto_pc<T>(q, sc, mixed);
    vx[k] = T(0);
}

// x_new = x + al ph + om sh; r = s - om t; partials |r|^2, <r0,r> re, im and |b - A x_new|^2
__global__ void hk_x(HOp A, const cxd* __restrict__ b, const cxd* __restrict__ x, cxd* xn, cxd* r,
                     const cxd* __restrict__ r0, const cxd* __restrict__ ph, const cxd* __restrict__ sh,
                     const cxd* __restrict__ sv, const cxd* __restrict__ t, const double* __restrict__ sc,
                     double* part, int nb)
{
    __shared__ double lds[8];
    HCELL;
    double a = 0.0, cre = 0.0, cim = 0.0, e = 0.0;
    if (i <= A.ni && j <= A.nj) {
        const size_t k = (size_t)i * A.ld + j;
        const cxd al = ldc(sc, HS_AL), om = ldc(sc, HS_OM);
        cxd xk(0.0);
        if (A.act[k]) {
            auto xnew = [&](size_t m) { return A.act[m] ? x[m] + al * ph[m] + om * sh[m] : cxd(0.0); };
            xk = xnew(k);
            const cxd q = sv[k] - om * t[k];
            r[k] = q;
            a = abs2v(q);
            const cxd c = conj(r0[k]) * q;
            cre = c.re;
            cim = c.im;
            const int ld = A.ld;
            const cxd rt = b[k] - (A.aP[k] * xk - A.aE[k] * xnew(k + ld) - A.aW[k] * xnew(k - ld) -
                                   A.aN[k] * xnew(k + 1) - A.aS[k] * xnew(k - 1));
            e = abs2v(rt);
        }
        xn[k] = xk;
    }
    const double t1 = block_sum<HX * HY>(a, lds);
    const double t2 = block_sum<HX * HY>(cre, lds);
    const double t3 = block_sum<HX * HY>(cim, lds);
    const double t4 = block_sum<HX * HY>(e, lds);
    if (lin_tid() == 0) {
        const int bid = blockIdx.y * gridDim.x + blockIdx.x;
        part[bid] = t1;
        part[nb + bid] = t2;
        part[2 * nb + bid] = t3;
        part[3 * nb + bid] = t4;
    }
}

__device__ inline double hsum(const double* part, int nb, double* lds)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) s += part[k];
    return block_sum<256>(s, lds);
}

// what: 1 restart (rr, rho1 = rr); 2 al = rho1/<r0,v>; 3 om = <t,s>/|t|^2; 4 after x
__global__ void hk_finish(const double* __restrict__ part, int nb, double* __restrict__ sc, int what)
{
    __shared__ double lds[8];
    const double a = hsum(part, nb, lds);
    const double b = hsum(part + nb, nb, lds);
    const double c = hsum(part + 2 * nb, nb, lds);
    const double d = what == 4 ? hsum(part + 3 * nb, nb, lds) : 0.0;
    if (threadIdx.x != 0) return;
    auto finite = [](cxd z) { return isfinite(z.re) && isfinite(z.im); };
    if (what == 1) {
        sc[HS_RR] = a;
        sc[HS_RT] = a;
        stc(sc, HS_RHO1, cxd(a, 0.0));
        stc(sc, HS_RHO, cxd(1.0));
        stc(sc, HS_AL, cxd(1.0));
        stc(sc, HS_OM, cxd(1.0));
    } else if (what == 2) {
        const cxd r0v(a, b);
        const cxd al = ldc(sc, HS_RHO1) / r0v;
        stc(sc, HS_AL, (abs2v(r0v) > 0.0 && finite(al)) ? al : cxd(0.0));
    } else if (what == 3) {
        const cxd ts(a, b);
        const cxd om = c > 0.0 ? ts / cxd(c) : cxd(0.0);
        stc(sc, HS_OM, finite(om) ? om : cxd(0.0));
    } else {
        sc[HS_RR] = a;
        stc(sc, HS_RHO, ldc(sc, HS_RHO1));
        stc(sc, HS_RHO1, cxd(b, c));
        sc[HS_RT] = d;
    }
}

// ------------------------------------------------------------------ solver
class HelmSolver {
public:
    ~HelmSolver()
    {
        if (h_sc) (void)hipHostFree(h_sc);
    }
    void configure(int ni, int nj);
    int ld0() const { return L64[0].ld; }
    cxd* aP() { return L64[0].aP; }
    cxd* aE() { return L64[0].aE; }
    cxd* aW() { return L64[0].aW; }
    cxd* aN() { return L64[0].aN; }
    cxd* aS() { return L64[0].aS; }
    cxd* rhs() { return bsol.p; }
    cxd* sol() { return xsol.p; }
    unsigned char* act() { return act0.p; }
    void solve(int maxITR, double tol, int mode, hipStream_t s, cathode_status* st);

private:
    void setup(int mixed, hipStream_t s);
    template <class T>
    void vcycle(Level<T>* L, hipStream_t s);
    HCoarsen cargs(int l);
    int nl = 0, lt = 0, cfg_ni = -1, cfg_nj = -1;
    Level<cxd> L64[HL_MAX];
    Level<cxf> L32[HL_MAX];
    size_t off[HL_MAX + 1] = {0};
    DevBuf<cxd> st64, mass, bsol, xsol, xalt, kv;
    DevBuf<cxf> st32;
    DevBuf<unsigned char> actall, act0;
    DevBuf<double> w0, part, sc;
    double* h_sc = nullptr;
};

static int hpad(int nj) { return ((nj + 2 + 7) / 8) * 8; }

void HelmSolver::configure(int ni, int nj)
{
    if (ni == cfg_ni && nj == cfg_nj) return;
    cfg_ni = ni;
    cfg_nj = nj;
    int d[HL_MAX][2];
    d[0][0] = ni;
    d[0][1] = nj;
    int n = 1;
    while (n < HL_MAX) {
        const int a = d[n - 1][0], b = d[n - 1][1];
        if ((long)a * b <= COARSEST_MAX || (a == 1 && b == 1)) break;
        d[n][0] = a > 1 ? (a + 1) / 2 : a;
        d[n][1] = b > 1 ? (b + 1) / 2 : b;
        n++;
    }
    nl = n;
    // the tail must fit 64 KiB of LDS at complex128 (7 values + flag per cell)
    lt = nl - 1;
    long cells = (long)(d[lt][0] + 2) * (d[lt][1] + 2);
    while (lt > 0) {
        const long c2 = cells + (long)(d[lt - 1][0] + 2) * (d[lt - 1][1] + 2);
        if ((long)d[lt - 1][0] * d[lt - 1][1] > TAIL_NT || c2 * (7 * 16 + 1) > 60000 || nl - lt >= TAIL_MAX_LEVELS) break;
        cells = c2;
        lt--;
    }
    size_t tot = 0;
    for (int l = 0; l < nl; l++) {
        off[l] = tot;
        tot += (size_t)(d[l][0] + 2) * hpad(d[l][1]);
    }
    off[nl] = tot;
    st64.alloc(8 * tot);
    st32.alloc(8 * tot);
    mass.alloc(tot);
    actall.alloc(tot);
    CATHODE_HIP(hipMemset(st64.p, 0, 8 * tot * sizeof(cxd)));
    CATHODE_HIP(hipMemset(st32.p, 0, 8 * tot * sizeof(cxf)));
    CATHODE_HIP(hipMemset(mass.p, 0, tot * sizeof(cxd)));
    CATHODE_HIP(hipMemset(actall.p, 0, tot));
    for (int l = 0; l < nl; l++) {
        const size_t o = off[l], sz = off[l + 1] - off[l];
        Level<cxd>& A = L64[l];
        Level<cxf>& B = L32[l];
        A.ni = B.ni = d[l][0];
        A.nj = B.nj = d[l][1];
        A.ld = B.ld = hpad(d[l][1]);
        A.cx = B.cx = (l + 1 < nl) && d[l][0] > 1;
        A.cy = B.cy = (l + 1 < nl) && d[l][1] > 1;
        cxd* p = st64.p;
        cxf* q = st32.p;
        A.aP = p + 5 * o; A.aE = A.aP + sz; A.aW = A.aE + sz; A.aN = A.aW + sz; A.aS = A.aN + sz;
        A.b = p + 5 * tot + o; A.x = p + 6 * tot + o; A.xu = p + 7 * tot + o;
        B.aP = q + 5 * o; B.aE = B.aP + sz; B.aW = B.aE + sz; B.aN = B.aW + sz; B.aS = B.aN + sz;
        B.b = q + 5 * tot + o; B.x = q + 6 * tot + o; B.xu = q + 7 * tot + o;
    }
    const size_t n0 = off[1];
    act0.alloc(n0);
    CATHODE_HIP(hipMemset(act0.p, 0, n0));
    bsol.alloc(n0);
    xsol.alloc(n0);
    xalt.alloc(n0);
    kv.alloc(8 * n0);
    CATHODE_HIP(hipMemset(bsol.p, 0, n0 * sizeof(cxd)));
    CATHODE_HIP(hipMemset(xsol.p, 0, n0 * sizeof(cxd)));
    CATHODE_HIP(hipMemset(xalt.p, 0, n0 * sizeof(cxd)));
    CATHODE_HIP(hipMemset(kv.p, 0, 8 * n0 * sizeof(cxd)));
    w0.alloc(nj + 2);
    part.alloc(4 * (cdiv(nj, HX) * cdiv(ni, HY)) + 64);
    sc.alloc(HS_COUNT);
    if (!h_sc) CATHODE_HIP(hipHostMalloc(&h_sc, HS_COUNT * sizeof(double), hipHostMallocDefault));
    // the plain hipMemset calls above run on the null stream, which the
    // library's non-blocking streams do not wait for
    CATHODE_HIP(hipDeviceSynchronize());
}

HCoarsen HelmSolver::cargs(int l)
{
    Level<cxd>& A = L64[l];
    Level<cxd>& C = L64[l + 1];
    Level<cxf>& G = L32[l + 1];
    HCoarsen a;
    a.fni = A.ni; a.fnj = A.nj; a.fld = A.ld; a.fi = A.cx ? 2 : 1; a.fj = A.cy ? 2 : 1;
    a.faP = A.aP; a.faE = A.aE; a.faW = A.aW; a.faN = A.aN; a.faS = A.aS;
    a.fm = mass.p + off[l];
    a.fact = actall.p + off[l];
    a.w = l == 0 ? w0.p : nullptr;
    a.cni = C.ni; a.cnj = C.nj; a.cld = C.ld;
    a.caP = C.aP; a.caE = C.aE; a.caW = C.aW; a.caN = C.aN; a.caS = C.aS; a.cb = C.b; a.cx = C.x;
    a.cm = mass.p + off[l + 1];
    a.cact = actall.p + off[l + 1];
    a.gP = G.aP; a.gE = G.aE; a.gW = G.aW; a.gN = G.aN; a.gS = G.aS;
    a.hx = A.cx ? 0.5 : 1.0;
    a.hy = A.cy ? 0.5 : 1.0;
    a.hb = (A.cx && A.cy) ? 0.5 : ((A.cx || A.cy) ? 0.75 : 1.0);
    return a;
}

void HelmSolver::setup(int mixed, hipStream_t s)
{
    Level<cxd>& F = L64[0];
    const int nb = cdiv(F.nj, HX) * cdiv(F.ni, HY);
    CATHODE_HIP(hipMemcpyAsync(actall.p, act0.p, off[1], hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(hk_prepare, hgrid(F.ni, F.nj), dim3(HX, HY), 0, s, F.ni, F.nj, F.ld, act0.p, F.aP, F.aE, F.aW,
                       F.aN, F.aS, bsol.p, F.b, part.p, nb);
    hipLaunchKernelGGL(hk_setup_finish, dim3(1), dim3(256), 0, s, part.p, nb, sc.p, (double)F.ni * F.nj);
    hipLaunchKernelGGL(hk_weights, dim3(1), dim3(256), (F.nj + 2) * sizeof(double), s, F.ni, F.nj, F.ld, F.aN, F.aS,
                       w0.p);
    Level<cxf>& G = L32[0];
    hipLaunchKernelGGL(hk_mass, hgrid(F.ni, F.nj), dim3(HX, HY), 0, s, F.ni, F.nj, F.ld, act0.p, F.b, w0.p, mass.p,
                       F.aP, F.aE, F.aW, F.aN, F.aS, G.aP, G.aE, G.aW, G.aN, G.aS, (const double*)sc.p, mixed);
    for (int l = 0; l + 1 < nl; l++) {
        Level<cxd>& C = L64[l + 1];
        hipLaunchKernelGGL(hk_coarsen, hgrid(C.ni, C.nj), dim3(HX, HY), 0, s, cargs(l), (const double*)sc.p, mixed);
    }
    CATHODE_LAUNCH_CHECK();
}

template <class T>
void HelmSolver::vcycle(Level<T>* L, hipStream_t s)
{
    for (int l = 0; l < lt; l++) {
        Level<T>& A = L[l];
        Level<T>& C = L[l + 1];
        VLevel<T> V{A.ni, A.nj, A.ld, A.aP, A.aE, A.aW, A.aN, A.aS, actall.p + off[l], A.b, A.x, nullptr};
        launch_vtile_cfg<T, true, HT_CFG>(l == 0, s, V, A.cx ? 2 : 1, A.cy ? 2 : 1, (const double*)w0.p, C.ni, C.nj,
                                           C.ld, C.b);
    }
    TailArgs<T> ta{};
    ta.nt = nl - lt;
    int o = 0;
    for (int q = 0; q < ta.nt; q++) {
        Level<T>& A = L[lt + q];
        ta.ni[q] = A.ni;
        ta.nj[q] = A.nj;
        ta.off[q] = o;
        o += (A.ni + 2) * (A.nj + 2);
        ta.cx[q] = A.cx;
        ta.cy[q] = A.cy;
        ta.gld[q] = A.ld;
        ta.aP[q] = A.aP;
        ta.aE[q] = A.aE;
        ta.aW[q] = A.aW;
        ta.aN[q] = A.aN;
        ta.aS[q] = A.aS;
        ta.act[q] = actall.p + off[lt + q];
    }
    ta.ncell = o;
    ta.b_in = L[lt].b;
    ta.x_out = L[lt].x;
    ta.w = lt == 0 ? w0.p : nullptr;
    hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(TAIL_NT), (size_t)o * (7 * sizeof(T) + 1) + 16, s, ta);
    for (int l = lt - 1; l >= 0; l--) {
        Level<T>& A = L[l];
        Level<T>& C = L[l + 1];
        VLevel<T> V{A.ni, A.nj, A.ld, A.aP, A.aE, A.aW, A.aN, A.aS, actall.p + off[l], A.b, A.x, A.xu};
        launch_vtile_cfg<T, false, HT_CFG>(l == 0, s, V, A.cx ? 2 : 1, A.cy ? 2 : 1, (const double*)nullptr, C.ni,
                                            C.nj, C.ld, l + 1 < lt ? C.xu : C.x);
    }
    CATHODE_LAUNCH_CHECK();
}

void HelmSolver::solve(int maxITR, double tol, int mode, hipStream_t s, cathode_status* st)
{
    auto t0 = std::chrono::steady_clock::now();
    Level<cxd>& F = L64[0];
    const size_t n0 = off[1];
    const int nb = cdiv(F.nj, HX) * cdiv(F.ni, HY);
    const int mixed = mode == CATHODE_MODE_MIXED ? 1 : 0;
    const dim3 g = hgrid(F.ni, F.nj), blk(HX, HY);
    setup(mixed, s);
    CATHODE_HIP(hipMemsetAsync(xsol.p, 0, n0 * sizeof(cxd), s));
    HOp A{F.ni, F.nj, F.ld, F.aP, F.aE, F.aW, F.aN, F.aS, act0.p};
    cxd *r = kv.p, *r0 = r + n0, *p = r0 + n0, *v = p + n0, *sv = v + n0, *t = sv + n0, *ph = t + n0, *sh = ph + n0;
    cxd *xa = xsol.p, *xb = xalt.p;
    // V-cycle output on level 0 (x when the whole hierarchy runs in the tail)
    const cxf* vout32 = lt > 0 ? L32[0].xu : L32[0].x;
    const cxd* vout64 = lt > 0 ? F.xu : F.x;
    int it = 0;
    double rel = 0.0;
    bool restart = true;
    while (it < maxITR) {
        if (restart) {
            hipLaunchKernelGGL(hk_restart, g, blk, 0, s, A, bsol.p, xa, r, r0, p, v, part.p, nb);
            hipLaunchKernelGGL(hk_finish, dim3(1), dim3(256), 0, s, part.p, nb, sc.p, 1);
        }
        it++;
        if (mixed) {
            hipLaunchKernelGGL(hk_p<cxf>, g, blk, 0, s, A, r, v, p, L32[0].b, L32[0].x, sc.p, restart ? 1 : 0, 1);
            vcycle<cxf>(L32, s);
            hipLaunchKernelGGL(hk_apply<cxf>, g, blk, 0, s, A, (const cxf*)vout32, ph, v, r0, sc.p, 1, 0, part.p, nb);
        } else {
            hipLaunchKernelGGL(hk_p<cxd>, g, blk, 0, s, A, r, v, p, F.b, F.x, sc.p, restart ? 1 : 0, 0);
            vcycle<cxd>(L64, s);
            hipLaunchKernelGGL(hk_apply<cxd>, g, blk, 0, s, A, (const cxd*)vout64, ph, v, r0, sc.p, 0, 0, part.p, nb);
        }
        hipLaunchKernelGGL(hk_finish, dim3(1), dim3(256), 0, s, part.p, nb, sc.p, 2);
        restart = false;
        if (mixed) {
            hipLaunchKernelGGL(hk_s<cxf>, g, blk, 0, s, A, r, v, sv, L32[0].b, L32[0].x, sc.p, 1);
            vcycle<cxf>(L32, s);
            hipLaunchKernelGGL(hk_apply<cxf>, g, blk, 0, s, A, (const cxf*)vout32, sh, t, sv, sc.p, 1, 1, part.p, nb);
        } else {
            hipLaunchKernelGGL(hk_s<cxd>, g, blk, 0, s, A, r, v, sv, F.b, F.x, sc.p, 0);
            vcycle<cxd>(L64, s);
            hipLaunchKernelGGL(hk_apply<cxd>, g, blk, 0, s, A, (const cxd*)vout64, sh, t, sv, sc.p, 0, 1, part.p, nb);
        }
        hipLaunchKernelGGL(hk_finish, dim3(1), dim3(256), 0, s, part.p, nb, sc.p, 3);
        hipLaunchKernelGGL(hk_x, g, blk, 0, s, A, bsol.p, xa, xb, r, r0, ph, sh, sv, t, sc.p, part.p, nb);
        hipLaunchKernelGGL(hk_finish, dim3(1), dim3(256), 0, s, part.p, nb, sc.p, 4);
        CATHODE_LAUNCH_CHECK();
        std::swap(xa, xb);
        fetch_host(h_sc, sc.p, HS_COUNT, s);
        CATHODE_HIP(hipStreamSynchronize(s));
        const double bb = h_sc[HS_BB];
        if (!(bb > 0.0)) {
            rel = bb == 0.0 ? 0.0 : NAN;
            if (bb == 0.0) it = 0;
            break;
        }
        rel = sqrt(h_sc[HS_RT] / bb);
        if (!std::isfinite(rel) || rel <= tol) break;
        if (sqrt(h_sc[HS_RR] / bb) <= tol) restart = true;
    }
    if (maxITR <= 0) {
        // no iteration: x = 0, relative residual 1 (0 when b = 0)
        fetch_host(h_sc, sc.p, HS_COUNT, s);
        CATHODE_HIP(hipStreamSynchronize(s));
        rel = h_sc[HS_BB] > 0.0 ? 1.0 : 0.0;
    }
    if (xa != xsol.p) CATHODE_HIP(hipMemcpyAsync(xsol.p, xa, n0 * sizeof(cxd), hipMemcpyDeviceToDevice, s));
    CATHODE_HIP(hipStreamSynchronize(s));
    if (st) {
        st->iterations = it;
        st->rel_residual = rel;
        st->converged = rel <= tol ? 1 : 0;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

// ------------------------------------------------------------------ E_theta assembly / post (B = 0)
constexpr double H_E0 = 1.60217662e-19, H_EPS0 = 8.8541878128e-12, H_MU0 = 1.2566370614e-6;

struct EthView {
    int ni, nj, ld, ld0;
    double dx, dr, rmin, omega, nu, epsr, masse;
    int pec_z, pec_rmax;
};

__device__ inline double eth_r(const EthView& V, int j) { return V.rmin + (j - 0.5) * V.dr; }

// sigma_theta-theta = n e^2 / (m (nu + i w) + EPS) / (1 + EPS), EPS = 1e-100 (1 + i)
__device__ inline cxd eth_sigma(const EthView& V, double ne)
{
    const cxd EPS(1e-100, 1e-100);
    const cxd nu_cmp(V.nu, V.omega);
    const cxd coef = cxd(ne * H_E0 * H_E0) / (cxd(V.masse) * nu_cmp + EPS) / (cxd(1.0) + cxd(0.0) + EPS);
    return coef * cxd(1.0);
}

__device__ inline cxd eth_R_open_z(const EthView& V)
{
    const double c0 = sqrt(1.0 / (H_EPS0 * V.epsr * H_MU0));
    const double deno = pow(2.0 * c0, 2) + pow(V.omega * V.dx, 2);
    cxd R(pow(2.0 * c0, 2) - pow(V.omega * V.dx, 2), -4.0 * c0 * V.omega * V.dx);
    return R / cxd(deno + 1e-100);
}

__device__ inline cxd eth_R_open_r(const EthView& V, int j)
{
    const double c0 = sqrt(1.0 / (H_EPS0 * V.epsr * H_MU0));
    const double r_t = (eth_r(V, j + 1) + eth_r(V, j)) / 2.0;
    const double qL = eth_r(V, j) / (r_t + 1e-100), qR = eth_r(V, j + 1) / (r_t + 1e-100);
    const double deno = pow(2.0 * c0 * sqrt(qR), 2) + pow(V.omega * V.dr * qR, 2);
    cxd R(pow(2.0 * c0, 2) * sqrt(qL * qR) - pow(V.omega * V.dr, 2) * qL * qR,
          -2.0 * c0 * V.omega * V.dr * (qL * sqrt(qR) + qR * sqrt(qL)));
    return R / cxd(deno + 1e-100);
}

// coefficients (emfieldModule.hpp:2414-2454) with the BC folds (:2713-2773)
__global__ void hk_eth_assemble(EthView V, const double* __restrict__ ne, const cxd* __restrict__ J, cxd* aP,
                                cxd* aE, cxd* aW, cxd* aN, cxd* aS, cxd* b)
{
#pragma clang fp contract(off)
    HCELL;
    if (i > V.ni || j > V.nj) return;
    const double dx = V.dx, dr = V.dr, w = V.omega;
    const double rj = eth_r(V, j), rm = eth_r(V, j - 1), rp = eth_r(V, j + 1);
    const double qL = (rm + rj) / 2.0 / (rj + 1e-100);
    const double qR = (rp + rj) / 2.0 / (rj + 1e-100);
    const size_t kf = (size_t)i * V.ld + j;
    const cxd sg = eth_sigma(V, ne[kf]);
    const cxd iwmu0(0.0, w * H_MU0);
    cxd P = cxd(2.0 * dx * dx + 2.0 * dr * dr + dx * dx * dr * dr / (rj * rj) -
                H_EPS0 * V.epsr * H_MU0 * w * w * dx * dx * dr * dr) +
            iwmu0 * sg * cxd(dx * dx * dr * dr);
    cxd E(dr * dr), W(dr * dr), N(qR * dx * dx), S(qL * dx * dx);
    const cxd bb = cxd(0.0, -w * H_MU0 * dx * dx * dr * dr) * J[kf];
    if (i == 1) {
        if (V.pec_z) P = P + W;
        else P = P - W * eth_R_open_z(V);
        W = cxd(0.0);
    }
    if (i == V.ni) {
        P = P + E;
        E = cxd(0.0);
    }
    if (j == 1 && V.rmin > 0.0) {
        P = P + cxd(rj / rm) * S;
        S = cxd(0.0);
    }
    if (j == V.nj) {
        if (V.pec_rmax) P = P + cxd(rj / rp) * N;
        else P = P - N * eth_R_open_r(V, j);
        N = cxd(0.0);
    }
    const size_t k = (size_t)i * V.ld0 + j;
    aP[k] = P;
    aE[k] = E;
    aW[k] = W;
    aN[k] = N;
    aS[k] = S;
    b[k] = bb;
}

// E with ghosts (emfieldModule.hpp:3213-3264), Pabs = 0.5 Re(J E*) with
// J = sigma E (:7733-7743); partial sums of Pabs 2 pi r dx dr
__global__ void hk_eth_post(EthView V, const double* __restrict__ ne, const cxd* __restrict__ sol, cxd* __restrict__ E,
                            double* __restrict__ Pabs, double* __restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double lds[8];
    const int j = blockIdx.x * HX + threadIdx.x;   // 0 .. nj+1
    const int i = blockIdx.y * HY + threadIdx.y;   // 0 .. ni+1
    double pw = 0.0;
    if (i <= V.ni + 1 && j <= V.nj + 1) {
        const bool in_i = i >= 1 && i <= V.ni, in_j = j >= 1 && j <= V.nj;
        const size_t kf = (size_t)i * V.ld + j;
        cxd e(0.0);
        if (in_i && in_j) {
            e = sol[(size_t)i * V.ld0 + j];
            const cxd Jp = eth_sigma(V, ne[kf]) * e;
            const double pa = 0.5 * (Jp * conj(e)).re;
            Pabs[kf] = pa;
            pw = pa * (V.dx * V.dr * 2.0 * M_PI * eth_r(V, j));
        } else if (i == 0 && in_j) {
            const cxd e1 = sol[(size_t)1 * V.ld0 + j];
            e = V.pec_z ? -e1 : eth_R_open_z(V) * e1;
        } else if (i == V.ni + 1 && in_j) {
            e = -sol[(size_t)V.ni * V.ld0 + j];
        } else if (j == 0 && in_i && V.rmin > 0.0) {
            e = cxd(-eth_r(V, 1) / eth_r(V, 0)) * sol[(size_t)i * V.ld0 + 1];
        } else if (j == V.nj + 1 && in_i) {
            const cxd en = sol[(size_t)i * V.ld0 + V.nj];
            e = V.pec_rmax ? cxd(-eth_r(V, V.nj) / eth_r(V, V.nj + 1)) * en : eth_R_open_r(V, V.nj) * en;
        }
        E[kf] = e;
    }
    const double t = block_sum<HX * HY>(pw, lds);
    if (lin_tid() == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = t;
}

static std::unique_ptr<HelmSolver> g_helm;
static DevBuf<cxd> g_hE, g_hJ;
static DevBuf<double> g_hne, g_hP, g_hpart;

void helm_release()
{
    g_helm.reset();
    g_hE.release();
    g_hJ.release();
    g_hne.release();
    g_hP.release();
    g_hpart.release();
}

}  // namespace cathode

using namespace cathode;

extern "C" {

int cathode_helm5_solve(const double* aP, const double* aE, const double* aW, const double* aN, const double* aS,
                        const double* b, const int* boxes, int n_bl, int ni2, int nj2, int ld, int maxITR, double tol,
                        int mode, double* x, cathode_status* st)
{
    CATHODE_GUARD_BEGIN
        if (!require_init()) return CATHODE_ERR_NOINIT;
        if (!aP || !aE || !aW || !aN || !aS || !b || !x || !boxes || n_bl <= 0 || ni2 < 3 || nj2 < 3 || ld < nj2 ||
            (mode != CATHODE_MODE_FP64 && mode != CATHODE_MODE_MIXED)) {
            set_error("helm5: invalid arguments");
            return CATHODE_ERR_ARG;
        }
        hipStream_t s = ctx().stream;
        if (!g_helm) g_helm.reset(new HelmSolver());
        HelmSolver& H = *g_helm;
        H.configure(ni2 - 2, nj2 - 2);
        const int ld0 = H.ld0();
        std::vector<unsigned char> mask((size_t)ni2 * ld0, 0);
        for (int q = 0; q < n_bl; q++) {
            const int* bx = boxes + 4 * q;
            if (bx[0] < 1 || bx[1] > ni2 - 2 || bx[2] < 1 || bx[3] > nj2 - 2) {
                set_error("helm5: boxes must lie inside the interior cells");
                return CATHODE_ERR_ARG;
            }
            for (int i = bx[0]; i <= bx[1]; i++)
                for (int j = bx[2]; j <= bx[3]; j++) mask[(size_t)i * ld0 + j] = 1;
        }
        CATHODE_HIP(hipMemcpyAsync(H.act(), mask.data(), mask.size(), hipMemcpyHostToDevice, s));
        const double* src[6] = {aP, aE, aW, aN, aS, b};
        cxd* dst[6] = {H.aP(), H.aE(), H.aW(), H.aN(), H.aS(), H.rhs()};
        for (int k = 0; k < 6; k++)
            CATHODE_HIP(hipMemcpy2DAsync(dst[k], ld0 * sizeof(cxd), src[k], ld * sizeof(cxd), nj2 * sizeof(cxd), ni2,
                                         hipMemcpyHostToDevice, s));
        cathode_status loc{};
        H.solve(maxITR, tol, mode, s, &loc);
        std::vector<cxd> hx((size_t)ni2 * ld0);
        CATHODE_HIP(hipMemcpyAsync(hx.data(), H.sol(), hx.size() * sizeof(cxd), hipMemcpyDeviceToHost, s));
        CATHODE_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < ni2; i++)
            for (int j = 0; j < nj2; j++)
                if (mask[(size_t)i * ld0 + j]) {
                    x[2 * ((size_t)i * ld + j)] = hx[(size_t)i * ld0 + j].re;
                    x[2 * ((size_t)i * ld + j) + 1] = hx[(size_t)i * ld0 + j].im;
                }
        record_status(loc);
        if (st) *st = loc;
        if (!std::isfinite(loc.rel_residual)) {
            set_error("helm5: non-finite residual");
            return CATHODE_ERR_NUMERIC;
        }
        return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_etheta_solve(const cathode_etheta_params* p, const double* ne, const double* J, double* E, double* Pabs,
                         double* P_total, cathode_status* st)
{
    CATHODE_GUARD_BEGIN
        if (!require_init()) return CATHODE_ERR_NOINIT;
        if (!p || !ne || !J || !E || p->ni < 2 || p->nj < 2 || !(p->dx > 0) || !(p->dr > 0) ||
            (p->mode != CATHODE_MODE_FP64 && p->mode != CATHODE_MODE_MIXED)) {
            set_error("etheta: invalid arguments");
            return CATHODE_ERR_ARG;
        }
        hipStream_t s = ctx().stream;
        if (!g_helm) g_helm.reset(new HelmSolver());
        HelmSolver& H = *g_helm;
        const int ni = p->ni, nj = p->nj, ld = nj + 2;
        const size_t n = (size_t)(ni + 2) * ld;
        H.configure(ni, nj);
        EthView V{ni, nj, ld, H.ld0(), p->dx, p->dr, p->rmin, p->omega, p->nu, p->epsr, p->masse, p->pec_z, p->pec_rmax};
        g_hne.alloc(n);
        g_hJ.alloc(n);
        g_hE.alloc(n);
        g_hP.alloc(n);
        const int nbp = cdiv(nj + 2, HX) * cdiv(ni + 2, HY);
        g_hpart.alloc(nbp + 64);
        CATHODE_HIP(hipMemcpyAsync(g_hne.p, ne, n * sizeof(double), hipMemcpyHostToDevice, s));
        CATHODE_HIP(hipMemcpyAsync(g_hJ.p, J, n * sizeof(cxd), hipMemcpyHostToDevice, s));
        std::vector<unsigned char> mask((size_t)(ni + 2) * H.ld0(), 0);
        for (int i = 1; i <= ni; i++)
            for (int j = 1; j <= nj; j++) mask[(size_t)i * H.ld0() + j] = 1;
        CATHODE_HIP(hipMemcpyAsync(H.act(), mask.data(), mask.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(hk_eth_assemble, hgrid(ni, nj), dim3(HX, HY), 0, s, V, g_hne.p, g_hJ.p, H.aP(), H.aE(), H.aW(),
                           H.aN(), H.aS(), H.rhs());
        CATHODE_LAUNCH_CHECK();
        cathode_status loc{};
        H.solve(p->maxITR, p->tol, p->mode, s, &loc);
        CATHODE_HIP(hipMemsetAsync(g_hP.p, 0, n * sizeof(double), s));
        hipLaunchKernelGGL(hk_eth_post, dim3(cdiv(nj + 2, HX), cdiv(ni + 2, HY)), dim3(HX, HY), 0, s, V, g_hne.p, H.sol(),
                           g_hE.p, g_hP.p, g_hpart.p);
        CATHODE_LAUNCH_CHECK();
        std::vector<double> part(nbp);
        CATHODE_HIP(hipMemcpyAsync(E, g_hE.p, n * sizeof(cxd), hipMemcpyDeviceToHost, s));
        if (Pabs) CATHODE_HIP(hipMemcpyAsync(Pabs, g_hP.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
        CATHODE_HIP(hipMemcpyAsync(part.data(), g_hpart.p, nbp * sizeof(double), hipMemcpyDeviceToHost, s));
        CATHODE_HIP(hipStreamSynchronize(s));
        double tot = 0.0;
        for (double v : part) tot += v;
        if (P_total) *P_total = tot;
        record_status(loc);
        if (st) *st = loc;
        if (!std::isfinite(loc.rel_residual)) {
            set_error("etheta: non-finite residual");
            return CATHODE_ERR_NUMERIC;
        }
        return CATHODE_OK;
    CATHODE_GUARD_END
}

}  // extern "C"
