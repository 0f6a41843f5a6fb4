// mg.hip -- multigrid kernels and the solve loop (see mg.hpp for the algorithm).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <chrono>

#include "mg.hpp"

namespace cathode {

// 2-D launch shape: x over j (contiguous in memory), y over i
constexpr int BX = 64, BY = 4;

static inline dim3 grid2(int ni, int nj) { return dim3(cdiv(nj, BX), cdiv(ni, BY)); }

#define CELL_IJ                                            \
    const int j = 1 + blockIdx.x * BX + threadIdx.x;       \
    const int i = 1 + blockIdx.y * BY + threadIdx.y;

// ------------------------------------------------------------------ level-0 setup
// Zero couplings to inactive neighbours (HYPRE's off-grid convention), make
// inactive cells identity rows, store the row excess s = max(aP - sum a, 0);
// block partials of max|aP| (slot 0) and |b|^2 (slot 1) over active cells.
__global__ void k_l0_prepare(int ni, int nj, int ld, const unsigned char* __restrict__ act,
                             double* aP, double* aE, double* aW, double* aN, double* aS,
                             const double* __restrict__ bsol, double* __restrict__ s_out,
                             double* __restrict__ part, int nb)
{
    __shared__ double lds[8];
    CELL_IJ;
    double amax = 0.0, bb = 0.0;
    if (i <= ni && j <= nj) {
        const size_t k = (size_t)i * ld + j;
        if (!act[k]) {
            aP[k] = 1.0;
            aE[k] = aW[k] = aN[k] = aS[k] = 0.0;
            s_out[k] = INFINITY;
        } else {
            double e = act[k + ld] ? aE[k] : 0.0;
            double w = act[k - ld] ? aW[k] : 0.0;
            double n = act[k + 1] ? aN[k] : 0.0;
            double s = act[k - 1] ? aS[k] : 0.0;
            aE[k] = e;
            aW[k] = w;
            aN[k] = n;
            aS[k] = s;
            double ex = aP[k] - (e + w + n + s);
            s_out[k] = ex > 0.0 ? ex : 0.0;
            amax = fabs(aP[k]);
            bb = bsol[k] * bsol[k];
        }
    }
    double v[2] = {amax, bb};
    const bool mx[2] = {true, false};
    block_reduce_vec<BX * BY, 2>(v, mx, lds);
    const double t1 = v[0], t2 = v[1];
    if (lin_tid() == 0) {
        const int bid = blockIdx.y * gridDim.x + blockIdx.x;
        part[bid] = t1;
        part[nb + bid] = t2;
    }
}

// Radial weights from column sums; one block.
__global__ void k_l0_weights(int ni, int nj, int ld, const double* __restrict__ aN,
                             const double* __restrict__ aS, double* __restrict__ w)
{
    extern __shared__ double ratio[];
    for (int j = 1 + threadIdx.x; j < nj; j += blockDim.x) {
        double sN = 0.0, sS = 0.0;
        for (int i = 1; i <= ni; i++) {
            sN += aN[(size_t)i * ld + j];
            sS += aS[(size_t)i * ld + j + 1];
        }
        double r = (sN > 0.0 && sS > 0.0) ? sN / sS : 1.0;
        ratio[j] = r < 1e-2 ? 1e-2 : (r > 1e2 ? 1e2 : r);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 1.0;
        w[0] = 1.0;
        w[1] = 1.0;
        for (int j = 1; j < nj; j++) {
            v *= ratio[j];
            w[j + 1] = v;
        }
        w[nj + 1] = v;
    }
}

__global__ void k_fill(double* p, int n, double v)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// one child of a coarse cell as the coarsening reads it
struct ChildVals {
    bool on;                     // present and active
    double p, e, w, n, s, m, wj;  // aP aE aW aN aS, mass, radial weight
};

// a level's block in LDS (k_setup_levels): cell (i, j) at (i - i0) * cols + (j - j0)
struct LevelStage {
    double *P, *E, *W, *N, *S, *M;
    unsigned char* act;
    int i0, j0, cols;
};

// coarse operator of coarse cell (I, J) from its (up to) 2 x 2 children in
// row-major order, fetch(i, j, here, ci0, cj0) giving each child's values
// (an absent child, here = false, reads child (ci0, cj0) and drops out, so
// every load can be issued before any is tested); writes the coarse level
// and, with stage_out, the values a further coarsening reads into out
// Returns the coarse row's excess fraction (aP - sum a_nb) / aP (+inf for an
// inactive coarse cell).
template <class Fetch>
__device__ inline double coarsen_cell_f(const CoarsenArgs& a, int I, int J, double s32, int mixed, Fetch fetch,
                                        const LevelStage& out, bool stage_out)
{
    const int i0 = a.fi * (I - 1) + 1, i1 = min(a.fi * I, a.fni);
    const int j0 = a.fj * (J - 1) + 1, j1 = min(a.fj * J, a.fnj);
    double S = 0.0, M = 0.0, E = 0.0, W = 0.0, N = 0.0, So = 0.0;
    int any = 0;
    ChildVals ch[4];
#pragma unroll
    for (int di = 0; di < 2; di++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++) {
            const int i = i0 + di, j = j0 + dj;
            ch[2 * di + dj] = fetch(i, j, i <= i1 && j <= j1, i0, j0);
        }
#pragma unroll
    for (int di = 0; di < 2; di++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++) {
            const int i = i0 + di, j = j0 + dj;
            const ChildVals& c = ch[2 * di + dj];
            if (!c.on) continue;
            any = 1;
            const double ex = c.p - (c.e + c.w + c.n + c.s);
            S += c.wj * (ex > 0.0 ? ex : 0.0);
            M += c.m;
            if (i == i1) E += c.wj * c.e;
            if (i == i0) W += c.wj * c.w;
            if (j == j1) N += c.wj * c.n;
            if (j == j0) So += c.wj * c.s;
        }
    const size_t kc = (size_t)I * a.cld + J;
    a.cact[kc] = (unsigned char)any;
    a.cb[kc] = 0.0;
    a.cx[kc] = 0.0;
    double P, frac = INFINITY;
    if (!any) {
        P = 1.0;
        E = W = N = So = 0.0;
        M = 0.0;
    } else {
        E *= a.hx;
        W *= a.hx;
        N *= a.hy;
        So *= a.hy;
        const double ex = M + a.hb * (S - M);
        P = ex + (E + W + N + So);
        frac = P > 0.0 ? ex / P : 0.0;
    }
    a.cm[kc] = M;
    a.caP[kc] = P;
    a.caE[kc] = E;
    a.caW[kc] = W;
    a.caN[kc] = N;
    a.caS[kc] = So;
    if (mixed) {
        a.gP[kc] = (float)(P * s32);
        a.gE[kc] = (float)(E * s32);
        a.gW[kc] = (float)(W * s32);
        a.gN[kc] = (float)(N * s32);
        a.gS[kc] = (float)(So * s32);
    }
    if (stage_out) {
        const int q = (I - out.i0) * out.cols + (J - out.j0);
        out.act[q] = (unsigned char)any;
        out.P[q] = P;
        out.E[q] = E;
        out.W[q] = W;
        out.N[q] = N;
        out.S[q] = So;
        out.M[q] = M;
    }
    return frac;
}

// a child read from the fine level in global memory
struct GlobalChild {
    const CoarsenArgs& a;
    __device__ ChildVals operator()(int i, int j, bool here, int ci0, int cj0) const
    {
        const size_t k = here ? (size_t)i * a.fld + j : (size_t)ci0 * a.fld + cj0;
        ChildVals c;
        c.on = here && a.fact[k] != 0;
        c.wj = a.w ? a.w[here ? j : cj0] : 1.0;
        c.p = a.faP[k];
        c.e = a.faE[k];
        c.w = a.faW[k];
        c.n = a.faN[k];
        c.s = a.faS[k];
        c.m = a.fm[k];
        return c;
    }
};

// the same from the fine level in global memory
__device__ inline double coarsen_cell(const CoarsenArgs& a, int I, int J, double s32, int mixed)
{
    return coarsen_cell_f(a, I, J, s32, mixed, GlobalChild{a}, LevelStage{}, false);
}

__global__ void k_coarsen(CoarsenArgs a, const double* __restrict__ sc, int mixed)
{
    const int J = 1 + blockIdx.x * BX + threadIdx.x;
    const int I = 1 + blockIdx.y * BY + threadIdx.y;
    if (I > a.cni || J > a.cnj) return;
    coarsen_cell(a, I, J, mixed ? sc[3] : 1.0, mixed);
}


__device__ inline double readlane_f64(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Dense inverse of the coarsest level (n = ni nj <= MG_COARSEST_MAX) by
// block Gauss-Jordan elimination without pivoting on the system padded to
// NM = MG_COARSEST_MAX unknowns (rows and columns n.. are the identity, so
// the leading n x n block of the result is A^-1), run by the NT / 64
// wavefronts of the workgroup.  Lane t holds column t of [A | I] (t < NM:
// A, t >= NM: I); wavefront w holds rows RPW w .. RPW w + RPW-1 of it in
// registers.  Block step K eliminates the RPW pivot columns of wavefront K's
// rows at once: that wavefront first reduces its own rows to the identity
// in those columns (Gauss-Jordan inside the wavefront, factors by
// readlane), publishes them in LDS (two buffers, so one barrier per block
// step), and every other row subtracts its entries in the pivot columns
// (read before the step, by readlane) times those rows.  With the fixed size
// every loop unrolls without branches.  The coarse operators are diagonally
// dominant with non-positive couplings (M-matrices), for which elimination
// without pivoting is stable.  A pivot at or below 1e-13 max|aP| (a singular
// operator, no Dirichlet part) zeroes its row, so the coarse solve returns a
// particular solution.
constexpr int CINV_NT = 512;
template <int NT>
__device__ __forceinline__ void coarse_inverse_block(const CoarseInv& c, double* lds_rows,
                                                     double* lds_red)
{
    constexpr int NM = MG_COARSEST_MAX;
    constexpr int RPW = NM / (NT / 64);  // rows per wavefront
    static_assert(NM % (NT / 64) == 0, "rows split evenly over the wavefronts");
    const int t = lin_tid() & 63, w = lin_tid() >> 6;
    const int n = c.ni * c.nj, nj = c.nj;
    double m[RPW];
    double amax = 0.0;
    // rows RPW w + r of column t: row q's diagonal, or its coupling to t when
    // t is q's neighbour (E of q is q + nj, W q - nj, N q + 1, S q - 1)
#pragma unroll
    for (int r = 0; r < RPW; r++) {
        const int q = w * RPW + r;
        double v = 0.0;
        if (t < NM) {
            if (q < n && t < n) {
                const int iq = q / nj, jq = q % nj;
                const size_t g = (size_t)(1 + iq) * c.ld + 1 + jq;
                const bool a = c.act[g] != 0;
                if (q == t) {
                    v = a ? c.aP[g] : 1.0;
                    amax = fabs(v);
                } else if (a) {
                    if (t == q + nj && iq + 1 < c.ni) v = -c.aE[g];
                    else if (t == q - nj && iq > 0) v = -c.aW[g];
                    else if (t == q + 1 && jq + 1 < nj) v = -c.aN[g];
                    else if (t == q - 1 && jq > 0) v = -c.aS[g];
                }
            } else if (q == t) {
                v = 1.0;  // padding
            }
        } else {
            v = (t - NM == q) ? 1.0 : 0.0;
        }
        m[r] = v;
    }
    const double tiny = 1e-13 * block_max_bcast<NT>(amax, lds_red);
#pragma unroll
    for (int K = 0; K < NT / 64; K++) {
        double* rows = lds_rows + (K & 1) * (RPW * 64);
        if (w == K) {
            // pivot columns RPW K .. RPW K + RPW-1 inside the wavefront
#pragma unroll
            for (int kk = 0; kk < RPW; kk++) {
                const int k = K * RPW + kk;
                const double pv = readlane_f64(m[kk], k);
                const double rcp = (fabs(pv) > tiny && isfinite(pv)) ? 1.0 / pv : 0.0;
                m[kk] = m[kk] * rcp;
#pragma unroll
                for (int r = 0; r < RPW; r++) {
                    if (r == kk) continue;
                    const double a = readlane_f64(m[r], k);
                    m[r] = fma(-a, m[kk], m[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < RPW; r++) rows[r * 64 + t] = m[r];
        }
        __syncthreads();
        if (w != K) {
            double f[RPW][RPW];
#pragma unroll
            for (int r = 0; r < RPW; r++)
#pragma unroll
                for (int kk = 0; kk < RPW; kk++) f[r][kk] = readlane_f64(m[r], K * RPW + kk);
#pragma unroll
            for (int kk = 0; kk < RPW; kk++) {
                const double p = rows[kk * 64 + t];
#pragma unroll
                for (int r = 0; r < RPW; r++) m[r] = fma(-f[r][kk], p, m[r]);
            }
        }
    }
    if (t >= NM && t - NM < n) {
        const int j = t - NM;
#pragma unroll
        for (int r = 0; r < RPW; r++) {
            const int q = w * RPW + r;
            if (q < n) c.inv64[q * n + j] = m[r];
        }
    }
}

// the chain's levels one after another, then (when ci.inv64 is set) the
// coarsest inverse; NT threads
// the chain's levels one after another, then (when ci.inv64 is set) the
// coarsest inverse; NT threads
template <int NT>
__device__ __forceinline__ void chain_block(const CoarsenChain& c, const double* __restrict__ sc, int mixed,
                                            const CoarseInv& ci)
{
    const double s32 = mixed ? sc[3] : 1.0;
    for (int l = 0; l < c.n; l++) {
        const CoarsenArgs& a = c.L[l];
        for (int q = lin_tid(); q < a.cni * a.cnj; q += NT) coarsen_cell(a, 1 + q / a.cnj, 1 + q % a.cnj, s32, mixed);
        __syncthreads();
    }
    if (ci.inv64) {
        __shared__ double rows[2 * (MG_COARSEST_MAX / (NT / 64)) * 64];
        __shared__ double red[NT / 64];
        coarse_inverse_block<NT>(ci, rows, red);
    }
}

__global__ void __launch_bounds__(512) k_coarsen_chain(CoarsenChain c, const double* __restrict__ sc, int mixed,
                                                      CoarseInv ci)
{
    chain_block<512>(c, sc, mixed, ci);
}

// the coarsest level's inverse when no chain builds it (it is a fine level)
__global__ void __launch_bounds__(CINV_NT) k_coarse_inverse(CoarseInv ci, const double* __restrict__ sc)
{
    __shared__ double rows[2 * (MG_COARSEST_MAX / (CINV_NT / 64)) * 64];
    __shared__ double red[CINV_NT / 64];
    (void)sc;
    coarse_inverse_block<CINV_NT>(ci, rows, red);
}

// A y_q at cell (i, j) of level 0 (5-point operator, couplings to inactive
// cells already zero); y_q has its own row pitch
__device__ inline double warm_ay(const WarmSet& ws, int q, int i, int j, size_t k, const double* aP,
                                 const double* aE, const double* aW, const double* aN, const double* aS)
{
    const double* y = ws.y[q];
    const int L = ws.ld[q];
    const size_t c = (size_t)i * L + j;
    return aP[k] * y[c] - (aE[k] * y[c + L] + aW[k] * y[c - L] + aN[k] * y[c + 1] + aS[k] * y[c - 1]);
}

// The setup after k_l0_prepare in one launch: the setup scalars (every
// workgroup sums the prepare partials in the same order, workgroup 0 stores
// them), the level-0 mass and FP32 copy, and the first nf coarsenings.
// Aggregates nest, so a workgroup owning an ri x rj block of level nf owns
// the cells of every finer level that block aggregates and coarsens them
// level by level with only its own barriers.
constexpr int SETUP_FUSE_MAX = 3;
constexpr int SETUP_NT = 256;
constexpr int SETUP_STAGE0 = 16 * 32;           // a workgroup's level-0 block
constexpr int SETUP_STAGE1 = SETUP_STAGE0 / 2;  // the next level's (at least one direction coarsens)
constexpr int SETUP_STAGE_COLS = 32;            // level-0 block columns (radial weights)
struct SetupLevels {
    int ni, nj, ld;
    const unsigned char* act;
    const double *s, *w;
    double* m;
    const double *aP, *aE, *aW, *aN, *aS;
    float *fP, *fE, *fW, *fN, *fS;
    const double* part;
    int nb;
    double nact;
    int nf, ri, rj;
    int pi[SETUP_FUSE_MAX + 1], pj[SETUP_FUSE_MAX + 1];  // level-l block = (ri pi[l]) x (rj pj[l]) cells
    CoarsenArgs L[SETUP_FUSE_MAX];
    // warm start of the coming solve: the dot products of its vectors, one
    // partial per workgroup (layout of WARM_NP, gridDim.x * gridDim.y long)
    WarmSet ws;
    const double* b;
    double* wpart;
    // the operator is the one of the previous setup (only b changed): the
    // scalars and the warm start's products, no level data
    int rhs_only;
    // V-cycle cut candidates: per-workgroup minima of the row excess
    // fraction of level 0 (dpart[bid]) and of level dcl (dpart[dnb + bid];
    // dcl in 1..nf, else not measured); dpart null: neither
    double* dpart;
    int dnb, dcl;
};

__global__ void __launch_bounds__(SETUP_NT) k_setup_levels(SetupLevels a, double* __restrict__ sc, int mixed)
{
    __shared__ double lds[SETUP_NT / 64];
    const int tid = threadIdx.x;
    double mx = 0.0, bb = 0.0;
    // two rounds of partial loads issued before they are combined (each sum
    // still runs over k = tid, tid + SETUP_NT, ... in order)
    for (int k = tid; k < a.nb; k += 2 * SETUP_NT) {
        const bool two = k + SETUP_NT < a.nb;
        const double m0 = a.part[k], b0 = a.part[a.nb + k];
        const double m1 = two ? a.part[k + SETUP_NT] : 0.0, b1 = two ? a.part[a.nb + k + SETUP_NT] : 0.0;
        mx = fmax(mx, m0);
        bb += b0;
        if (two) {
            mx = fmax(mx, m1);
            bb += b1;
        }
    }
    mx = block_max_bcast<SETUP_NT>(mx, lds);
    bb = block_sum_bcast<SETUP_NT>(bb, lds);
    int e = 0;
    (void)frexp(mx > 0.0 ? mx : 1.0, &e);
    const double s32 = ldexp(1.0, -e);  // s32 * max|aP| in [0.5, 1)
    if (tid == 0 && blockIdx.x == 0 && blockIdx.y == 0) {
        sc[0] = bb;                                    // |b|^2
        sc[2] = bb > 0.0 ? sqrt(a.nact / bb) : 1.0;   // mixed input scale
        sc[3] = s32;
    }
    // Each level of this workgroup's blocks is also staged in LDS (act, aP..aS,
    // mass; level l in buffer set l & 1), so every coarsening reads its
    // children there instead of waiting for the global stores of the level
    // before.  Staged only when the blocks fit the buffers (always, for the
    // 16 x 32 level-0 block of every fused depth).
    __shared__ double st0[6 * SETUP_STAGE0], st1[6 * SETUP_STAGE1];
    __shared__ unsigned char sa0[SETUP_STAGE0], sa1[SETUP_STAGE1];
    __shared__ double wst[SETUP_STAGE_COLS];
    auto stage = [&](int l, int i0, int j0, int cols) {
        double* b = (l & 1) ? st1 : st0;
        const int n = (l & 1) ? SETUP_STAGE1 : SETUP_STAGE0;
        return LevelStage{b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n, (l & 1) ? sa1 : sa0, i0, j0, cols};
    };
    if (a.rhs_only) {
        if (a.ws.k == 0) return;
        double wd[WARM_NP];
#pragma unroll
        for (int q = 0; q < WARM_NP; q++) wd[q] = 0.0;
        const int rows = a.ri * a.pi[0], cols = a.rj * a.pj[0];
        const int i0 = 1 + (int)blockIdx.y * rows, j0 = 1 + (int)blockIdx.x * cols;
        for (int q = tid; q < rows * cols; q += SETUP_NT) {
            const int i = i0 + q / cols, j = j0 + q % cols;
            if (i > a.ni || j > a.nj) continue;
            const size_t k = (size_t)i * a.ld + j;
            if (!a.act[k]) continue;
            double ay[WARM_MAX];
#pragma unroll
            for (int q2 = 0; q2 < WARM_MAX; q2++)
                ay[q2] = q2 < a.ws.k ? warm_ay(a.ws, q2, i, j, k, a.aP, a.aE, a.aW, a.aN, a.aS) : 0.0;
            const double bk = a.b[k];
#pragma unroll
            for (int q2 = 0; q2 < WARM_MAX; q2++) {
                wd[q2] += ay[q2] * bk;
#pragma unroll
                for (int p = 0; p <= q2; p++) wd[WARM_MAX + warm_gidx(p, q2)] += ay[p] * ay[q2];
            }
            wd[WARM_NP - 1] += bk * bk;
        }
        __shared__ double wl0[WARM_NP * (SETUP_NT / 64)];
        block_sum_vec_bcast<SETUP_NT, WARM_NP>(wd, wl0);
        if (tid == 0) {
            const int nsb = (int)(gridDim.x * gridDim.y), bid = (int)(blockIdx.y * gridDim.x + blockIdx.x);
#pragma unroll
            for (int q = 0; q < WARM_NP; q++) a.wpart[q * nsb + bid] = wd[q];
        }
        return;
    }
    bool staged = a.rj * a.pj[0] <= SETUP_STAGE_COLS;
    for (int l = 0; l <= a.nf; l++)
        staged = staged && a.ri * a.pi[l] * a.rj * a.pj[l] <= ((l & 1) ? SETUP_STAGE1 : SETUP_STAGE0);
    // level 0: mass part m = w_j min(s) over the 5-point neighbourhood, FP32
    // copy, and the warm start's dot products
    double wd[WARM_NP];
#pragma unroll
    for (int q = 0; q < WARM_NP; q++) wd[q] = 0.0;
    double d0 = INFINITY, dc = INFINITY;  // excess fraction minima of level 0 and level dcl
    {
        const int rows = a.ri * a.pi[0], cols = a.rj * a.pj[0];
        const int i0 = 1 + (int)blockIdx.y * rows, j0 = 1 + (int)blockIdx.x * cols;
        const LevelStage S0 = stage(0, i0, j0, cols);
        if (staged && tid < cols) wst[tid] = j0 + tid <= a.nj + 1 ? a.w[j0 + tid] : 1.0;
        for (int q = tid; q < rows * cols; q += SETUP_NT) {
            const int i = i0 + q / cols, j = j0 + q % cols;
            if (staged) S0.act[q] = 0;
            if (i > a.ni || j > a.nj) continue;
            const size_t k = (size_t)i * a.ld + j;
            if (mixed) {
                a.fP[k] = (float)(a.aP[k] * s32);
                a.fE[k] = (float)(a.aE[k] * s32);
                a.fW[k] = (float)(a.aW[k] * s32);
                a.fN[k] = (float)(a.aN[k] * s32);
                a.fS[k] = (float)(a.aS[k] * s32);
            }
            if (!a.act[k]) {
                a.m[k] = 0.0;
                continue;
            }
            if (a.ws.k > 0) {
                double ay[WARM_MAX];
#pragma unroll
                for (int q = 0; q < WARM_MAX; q++)
                    ay[q] = q < a.ws.k ? warm_ay(a.ws, q, i, j, k, a.aP, a.aE, a.aW, a.aN, a.aS) : 0.0;
                const double bk = a.b[k];
#pragma unroll
                for (int q = 0; q < WARM_MAX; q++) {
                    wd[q] += ay[q] * bk;
#pragma unroll
                    for (int p = 0; p <= q; p++) wd[WARM_MAX + warm_gidx(p, q)] += ay[p] * ay[q];
                }
                wd[WARM_NP - 1] += bk * bk;
            }
            d0 = fmin(d0, a.aP[k] > 0.0 ? a.s[k] / a.aP[k] : 0.0);
            double v = a.s[k];
            if (a.act[k + a.ld]) v = fmin(v, a.s[k + a.ld]);
            if (a.act[k - a.ld]) v = fmin(v, a.s[k - a.ld]);
            if (a.act[k + 1]) v = fmin(v, a.s[k + 1]);
            if (a.act[k - 1]) v = fmin(v, a.s[k - 1]);
            const double mk = a.w[j] * v;
            a.m[k] = mk;
            if (staged) {
                S0.act[q] = 1;
                S0.P[q] = a.aP[k];
                S0.E[q] = a.aE[k];
                S0.W[q] = a.aW[k];
                S0.N[q] = a.aN[k];
                S0.S[q] = a.aS[k];
                S0.M[q] = mk;
            }
        }
    }
    if (a.ws.k > 0) {
        __shared__ double wl[WARM_NP * (SETUP_NT / 64)];
        block_sum_vec_bcast<SETUP_NT, WARM_NP>(wd, wl);
        if (tid == 0) {
            const int nsb = (int)(gridDim.x * gridDim.y), bid = (int)(blockIdx.y * gridDim.x + blockIdx.x);
#pragma unroll
            for (int q = 0; q < WARM_NP; q++) a.wpart[q * nsb + bid] = wd[q];
        }
    }
    for (int l = 0; l < a.nf; l++) {
        __syncthreads();  // level l of this block staged (LDS) and written (global, same workgroup)
        const CoarsenArgs& c = a.L[l];
        const int fr = a.ri * a.pi[l], fc = a.rj * a.pj[l];
        const int rows = a.ri * a.pi[l + 1], cols = a.rj * a.pj[l + 1];
        const int I0 = 1 + (int)blockIdx.y * rows, J0 = 1 + (int)blockIdx.x * cols;
        const LevelStage src = stage(l, 1 + (int)blockIdx.y * fr, 1 + (int)blockIdx.x * fc, fc);
        const LevelStage dst = stage(l + 1, I0, J0, cols);
        for (int q = tid; q < rows * cols; q += SETUP_NT) {
            const int I = I0 + q / cols, J = J0 + q % cols;
            if (staged) dst.act[q] = 0;
            if (I > c.cni || J > c.cnj) continue;
            if (!staged) {
                const double f = coarsen_cell(c, I, J, s32, mixed);
                if (l + 1 == a.dcl) dc = fmin(dc, f);
                continue;
            }
            const double f = coarsen_cell_f(
                c, I, J, s32, mixed,
                [&](int i, int j, bool here, int ci0, int cj0) {
                    const int ii = here ? i : ci0, jj = here ? j : cj0;
                    const int lq = (ii - src.i0) * src.cols + (jj - src.j0);
                    ChildVals v;
                    v.on = here && src.act[lq] != 0;
                    v.wj = c.w ? wst[jj - src.j0] : 1.0;  // the weights apply to level 0 only
                    v.p = src.P[lq];
                    v.e = src.E[lq];
                    v.w = src.W[lq];
                    v.n = src.N[lq];
                    v.s = src.S[lq];
                    v.m = src.M[lq];
                    return v;
                },
                dst, true);
            if (l + 1 == a.dcl) dc = fmin(dc, f);
        }
    }
    if (a.dpart) {
        __syncthreads();  // lds is free again
        d0 = -block_max<SETUP_NT>(-d0, lds);
        dc = -block_max<SETUP_NT>(-dc, lds);
        if (tid == 0) {
            const int bid = (int)(blockIdx.y * gridDim.x + blockIdx.x);
            a.dpart[bid] = d0;
            a.dpart[a.dnb + bid] = dc;
        }
    }
}

// ------------------------------------------------------------------ V-cycle kernels
__device__ inline double abs2v(double v) { return v * v; }
__device__ inline float abs2v(float v) { return v * v; }
}  // namespace cathode
#include "vcycle.hpp"
namespace cathode {

// ------------------------------------------------------------------ outer loop kernels
// Device scalars (MGSolver::scal):
enum {
    SC_BB = 0,    // |b|^2
    SC_RR = 1,    // |r|^2 (recursive residual)
    SC_SCALE = 2, // mixed: preconditioner input scale sqrt(n)/|b|
    SC_S32 = 3,   // mixed: power-of-two matrix scale of the FP32 levels
    SC_RHO = 4,   // BiCGStab rho_{k-1}
    SC_RHO1 = 5,  // BiCGStab rho_k = (r0, r)
    SC_AL = 6,    // alpha
    SC_OM = 7,    // omega
    SC_NC = 8,    // (int) coarsest unknowns
    SC_R0V = 9,
    SC_TS = 10,
    SC_TT = 11,
    SC_RT = 12,   // |b - A x|^2 (true residual)
    SC_D0 = 13,   // V-cycle cut candidates: row excess fraction minimum of level 0
    SC_DC = 14,   // ... and of the second candidate level
    SC_SIDE = 15, // caller's value read with the first round trip (see MGSolver::side_slot)
    SC_COUNT = 16
};

// sum of squares of the active entries of v (partials per block)
__global__ void k_sumsq(int ni, int nj, int ld, const double* __restrict__ v,
                        const unsigned char* __restrict__ act, double* __restrict__ part)
{
    __shared__ double lds[8];
    CELL_IJ;
    double s = 0.0;
    if (i <= ni && j <= nj) {
        const size_t k = (size_t)i * ld + j;
        if (act[k]) s = v[k] * v[k];
    }
    double tot = block_sum<BX * BY>(s, lds);
    if (lin_tid() == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = tot;
}

// preconditioner input: vin = T(src * in_scale), vx = 0
template <class T>
__device__ inline T to_pc(double v, const double* sc, int mixed)
{
    return mixed ? (T)(v * sc[SC_S32] * sc[SC_SCALE]) : (T)v;
}
template <class T>
__device__ inline double from_pc(T y, const double* sc, int mixed)
{
    return mixed ? (double)y / sc[SC_SCALE] : (double)y;
}

// Sum over the off-diagonal couplings of coupling * neighbour value: the
// level-0 5-point arrays, or (nine != nullptr) the nine-point couplings
// E W N S EN ES WN WS stored n0 apart.  xat(m) gives the neighbour value;
// inactive and ghost cells hold 0 in every vector it is used on.
template <class F>
__device__ inline double offdiag(const double* __restrict__ aE, const double* __restrict__ aW,
                                 const double* __restrict__ aN, const double* __restrict__ aS,
                                 const double* __restrict__ nine, size_t n0, size_t k, int ld, F xat)
{
    if (!nine) return aE[k] * xat(k + ld) + aW[k] * xat(k - ld) + aN[k] * xat(k + 1) + aS[k] * xat(k - 1);
    return nine[k] * xat(k + ld) + nine[n0 + k] * xat(k - ld) + nine[2 * n0 + k] * xat(k + 1) +
           nine[3 * n0 + k] * xat(k - 1) + nine[4 * n0 + k] * xat(k + ld + 1) +
           nine[5 * n0 + k] * xat(k + ld - 1) + nine[6 * n0 + k] * xat(k - ld + 1) +
           nine[7 * n0 + k] * xat(k - ld - 1);
}

// level-0 5-point preconditioner couplings from the nine-point ones
__global__ void k_lump9(int ni, int nj, int ld, const double* __restrict__ nine, size_t n0,
                        const unsigned char* __restrict__ act, double* __restrict__ aE, double* __restrict__ aW,
                        double* __restrict__ aN, double* __restrict__ aS)
{
    CELL_IJ;
    if (i > ni || j > nj) return;
    const size_t k = (size_t)i * ld + j;
    const double en = act[k + ld + 1] ? nine[4 * n0 + k] : 0.0, es = act[k + ld - 1] ? nine[5 * n0 + k] : 0.0;
    const double wn = act[k - ld + 1] ? nine[6 * n0 + k] : 0.0, ws = act[k - ld - 1] ? nine[7 * n0 + k] : 0.0;
    aE[k] = nine[k] + 0.5 * (en + es);
    aW[k] = nine[n0 + k] + 0.5 * (wn + ws);
    aN[k] = nine[2 * n0 + k] + 0.5 * (en + wn);
    aS[k] = nine[3 * n0 + k] + 0.5 * (es + ws);
}

// r = b - A x, r0 = r, p = v = 0 (restart); partials |r|^2 (slot 0), (r0,r) (slot 1)

// the least-squares coefficients of the first n vectors from the summed
// dot products t: Gaussian elimination on the (symmetric positive
// semidefinite) Gram matrix without pivoting; a pivot at or below 1e-12
// times its diagonal entry (a vector in the span of the earlier ones) drops
// that vector.  Returns |b - A x0|^2 = <b, b> - sum a_q <Ay_q, b>.
__device__ inline double warm_coeffs(const double* t, int n, double* a)
{
    double G[WARM_MAX][WARM_MAX], r[WARM_MAX], d0[WARM_MAX];
    bool on[WARM_MAX];
#pragma unroll
    for (int p = 0; p < WARM_MAX; p++) {
        r[p] = p < n ? t[p] : 0.0;
#pragma unroll
        for (int q = 0; q < WARM_MAX; q++)
            G[p][q] = (p < n && q < n) ? t[WARM_MAX + (p <= q ? warm_gidx(p, q) : warm_gidx(q, p))]
                                       : (p == q ? 1.0 : 0.0);
        d0[p] = G[p][p];
        on[p] = p < n && d0[p] > 0.0;
    }
#pragma unroll
    for (int p = 0; p < WARM_MAX; p++) {
        if (on[p] && !(G[p][p] > 1e-12 * d0[p])) on[p] = false;
        if (!on[p]) {
#pragma unroll
            for (int q = 0; q < WARM_MAX; q++) G[p][q] = G[q][p] = 0.0;
            G[p][p] = 1.0;
            r[p] = 0.0;
            continue;
        }
#pragma unroll
        for (int q = p + 1; q < WARM_MAX; q++) {
            const double f = G[q][p] / G[p][p];
#pragma unroll
            for (int c = p; c < WARM_MAX; c++) G[q][c] -= f * G[p][c];
            r[q] -= f * r[p];
        }
    }
#pragma unroll
    for (int p = WARM_MAX - 1; p >= 0; p--) {
        double s = r[p];
#pragma unroll
        for (int q = p + 1; q < WARM_MAX; q++) s -= G[p][q] * a[q];
        a[p] = on[p] ? s / G[p][p] : 0.0;
    }
    double red = 0.0;
    bool ok = true;
#pragma unroll
    for (int p = 0; p < WARM_MAX; p++) {
        ok = ok && isfinite(a[p]);
        red += p < n ? a[p] * t[p] : 0.0;
    }
    if (!ok) {
#pragma unroll
        for (int p = 0; p < WARM_MAX; p++) a[p] = 0.0;
        red = 0.0;
    }
    return t[WARM_NP - 1] - red;
}

struct RestartArgs {
    int ni, nj, ld;
    const double *aP, *aE, *aW, *aN, *aS, *nine;
    size_t n0;
    const double *b, *x;
    const unsigned char* act;
    double *r, *r0, *p, *v, *part;
    int nb;
    double* xout;
    WarmSet ws;
    const double* wpart;
    int wnb;
    // the setup launch's excess fraction partials (nullptr: none), reduced
    // by restart block 0 into sc[SC_D0], sc[SC_DC]
    const double* dpart;
    int dnb, dstride;  // workgroups of the setup launch, distance of the second set
    double* sc;
};

// NT / (BX BY) blocks of the restart in one workgroup of BX x (NT / BX)
// threads: blocks blk0, blk0 + 1, ... of the 2-D block grid in row-major
// order, one per BX x BY threads.  Every block's partial sums its four
// wavefronts in the same order whatever NT is.
template <int NT>
__device__ __forceinline__ void restart_block(const RestartArgs& A, int blk0)
{
    __shared__ double lds[NT / 64];
    __shared__ double wl[WARM_NP * (NT / 64)];
    const int ni = A.ni, nj = A.nj, ld = A.ld;
    const double *aP = A.aP, *aE = A.aE, *aW = A.aW, *aN = A.aN, *aS = A.aS, *nine = A.nine, *b = A.b, *x = A.x;
    const size_t n0 = A.n0;
    const unsigned char* act = A.act;
    const WarmSet& ws = A.ws;
    const int nbx = (nj + BX - 1) / BX, nby = (ni + BY - 1) / BY;
    const int bid = blk0 + lin_tid() / (BX * BY);
    const bool live = bid < nbx * nby;
    const int j = 1 + (bid % nbx) * BX + (int)threadIdx.x;
    const int i = 1 + (bid / nbx) * BY + (int)threadIdx.y % BY;
    // warm start x = sum a_q y_q from the caller's partials (every workgroup
    // sums them in the same order and solves the same small system)
    double a[WARM_MAX] = {0.0, 0.0, 0.0, 0.0};
    if (ws.k > 0) {
        double t[WARM_NP];
#pragma unroll
        for (int m = 0; m < WARM_NP; m++) t[m] = 0.0;
        // all WARM_NP loads of a partial index issued together (each sum
        // still runs over q = t, t + NT, ... in order)
        for (int q = lin_tid(); q < A.wnb; q += NT) {
            double v[WARM_NP];
#pragma unroll
            for (int m = 0; m < WARM_NP; m++) v[m] = A.wpart[m * A.wnb + q];
#pragma unroll
            for (int m = 0; m < WARM_NP; m++) t[m] += v[m];
        }
        block_sum_vec_bcast<NT, WARM_NP>(t, wl);
        const double res = warm_coeffs(t, ws.k, a);
        if (ws.report && lin_tid() == 0 && blk0 == 0) {
            double a2[WARM_MAX];
            const double res_less = warm_coeffs(t, ws.k - 1, a2);
            ws.report[0] = res_less > 0.0 ? fmax(res, 0.0) / res_less : 1.0;
        }
    }
    double rr = 0.0;
    if (live && i <= ni && j <= nj) {
        const size_t k = (size_t)i * ld + j;
        double q = 0.0;
        if (ws.k > 0) {
            // x = sum a_q y_q on the interior (the y_q have zero ghosts), r = b - A x
            double x0 = 0.0, ax = 0.0;
#pragma unroll
            for (int m = 0; m < WARM_MAX; m++)
                if (m < ws.k) {
                    x0 += a[m] * ws.y[m][(size_t)i * ws.ld[m] + j];
                    if (act[k]) ax += a[m] * warm_ay(ws, m, i, j, k, aP, aE, aW, aN, aS);
                }
            A.xout[k] = x0;
            if (act[k]) {
                q = b[k] - ax;
                rr = q * q;
            }
        } else if (A.xout) {
            // zero initial guess: x = 0 on the interior, r = b
            A.xout[k] = 0.0;
            if (act[k]) {
                q = b[k];
                rr = q * q;
            }
        } else if (act[k]) {
            q = b[k] - aP[k] * x[k] + offdiag(aE, aW, aN, aS, nine, n0, k, ld, [&](size_t m) { return x[m]; });
            rr = q * q;
        }
        A.r[k] = q;
        A.r0[k] = q;
        A.p[k] = 0.0;
        A.v[k] = 0.0;
    }
    if (blk0 == 0 && A.dpart) {
        double d0 = INFINITY, dc = INFINITY;
        for (int q = lin_tid(); q < A.dnb; q += NT) {
            d0 = fmin(d0, A.dpart[q]);
            dc = fmin(dc, A.dpart[A.dstride + q]);
        }
        d0 = -block_max<NT>(-d0, lds);
        dc = -block_max<NT>(-dc, lds);
        if (lin_tid() == 0) {
            A.sc[SC_D0] = d0;
            A.sc[SC_DC] = dc;
        }
    }
    // the k_bicg_x layout: (r, r), (r0, r) = (r, r), |b - A x|^2 = (r, r)
    rr = wave_sum(rr);
    const int w = lin_tid() >> 6;
    if ((lin_tid() & 63) == 0) lds[w] = rr;
    __syncthreads();
    if (live && lin_tid() % (BX * BY) == 0) {
        double t = 0.0;
        for (int q = 0; q < BX * BY / 64; q++) t += lds[w + q];
        A.part[bid] = t;
        A.part[A.nb + bid] = t;
        A.part[2 * A.nb + bid] = t;
    }
}

__global__ void __launch_bounds__(BX * BY) k_bicg_restart(RestartArgs A)
{
    restart_block<BX * BY>(A, (int)(blockIdx.y * gridDim.x + blockIdx.x));
}

// The first restart of a solve together with the coarse chain of its setup:
// both read only what the setup launch wrote and write disjoint buffers.
// Workgroup 0 (dispatched first) builds the chain's levels and the coarsest
// inverse; workgroups 1.. take two restart blocks each, in row-major order.
constexpr int RC_NT = 2 * BX * BY;
__global__ void __launch_bounds__(RC_NT) k_restart_chain(RestartArgs A, CoarsenChain c, const double* __restrict__ sc,
                                                        int mixed, CoarseInv ci)
{
    if (blockIdx.x == 0) {
        chain_block<RC_NT>(c, sc, mixed, ci);
        return;
    }
    restart_block<RC_NT>(A, 2 * ((int)blockIdx.x - 1));
}

// p = r + beta (p - om v), beta = (rho1/rho)(al/om) (0 after a restart);
// preconditioner input vin = p, vx = 0
template <class T>
__global__ void k_bicg_p(int ni, int nj, int ld, const unsigned char* __restrict__ act,
                         const double* __restrict__ r, const double* __restrict__ v,
                         double* __restrict__ p, T* __restrict__ vin, T* __restrict__ vx,
                         const double* __restrict__ sc, int first, int mixed)
{
    CELL_IJ;
    if (i > ni || j > nj) return;
    const size_t k = (size_t)i * ld + j;
    double pk = 0.0;
    if (act[k]) {
        if (first) {
            pk = r[k];
        } else {
            const double beta = bicg_beta(sc[SC_RHO1], sc[SC_RHO], sc[SC_AL], sc[SC_OM]);
            pk = r[k] + beta * (p[k] - sc[SC_OM] * v[k]);
        }
    }
    p[k] = pk;
    vin[k] = to_pc<T>(pk, sc, mixed);
    vx[k] = 0;
}

// y = M^-1(in) from the V-cycle output; out = A y; partial dot (dotvec, out)
// -> slot 0; with want_tt also (out, out) -> slot 1.
template <class T>
__global__ void k_bicg_apply(int ni, int nj, int ld, const double* __restrict__ aP,
                             const double* __restrict__ aE, const double* __restrict__ aW,
                             const double* __restrict__ aN, const double* __restrict__ aS,
                             const double* __restrict__ nine, size_t n0,
                             const unsigned char* __restrict__ act, const T* __restrict__ vx,
                             double* __restrict__ y, double* __restrict__ out,
                             const double* __restrict__ dotvec, const double* __restrict__ sc, int mixed,
                             int want_tt, double* __restrict__ part, int nb)
{
    __shared__ double lds[8];
    CELL_IJ;
    double d1 = 0.0, d2 = 0.0;
    if (i <= ni && j <= nj) {
        const size_t k = (size_t)i * ld + j;
        double o = 0.0;
        double yk = 0.0;
        if (act[k]) {
            // inactive neighbours have zero coupling and a zero V-cycle output
            yk = from_pc<T>(vx[k], sc, mixed);
            o = aP[k] * yk -
                offdiag(aE, aW, aN, aS, nine, n0, k, ld, [&](size_t m) { return from_pc<T>(vx[m], sc, mixed); });
            d1 = dotvec[k] * o;
            d2 = o * o;
        }
        y[k] = yk;
        out[k] = o;
    }
    double t1 = block_sum<BX * BY>(d1, lds);
    double t2 = want_tt ? block_sum<BX * BY>(d2, lds) : 0.0;
    if (lin_tid() == 0) {
        const int bid = blockIdx.y * gridDim.x + blockIdx.x;
        part[bid] = t1;
        if (want_tt) part[nb + bid] = t2;
    }
}

// s = r - al v, preconditioner input vin = s, vx = 0
template <class T>
__global__ void k_bicg_s(int ni, int nj, int ld, const unsigned char* __restrict__ act,
                         const double* __restrict__ r, const double* __restrict__ v,
                         double* __restrict__ sv, T* __restrict__ vin, T* __restrict__ vx,
                         const double* __restrict__ sc, int mixed)
{
    CELL_IJ;
    if (i > ni || j > nj) return;
    const size_t k = (size_t)i * ld + j;
    const double q = act[k] ? r[k] - sc[SC_AL] * v[k] : 0.0;
    sv[k] = q;
    vin[k] = to_pc<T>(q, sc, mixed);
    vx[k] = 0;
}

// x_new = x + al ph + om sh (into the other x buffer); r = s - om t;
// partials |r|^2 (slot 0), (r0, r) (slot 1) and the true residual
// |b - A x_new|^2 (slot 2), evaluating the neighbours' x_new from the old buffer
__global__ void k_bicg_x(int ni, int nj, int ld, const double* __restrict__ aP, const double* __restrict__ aE,
                         const double* __restrict__ aW, const double* __restrict__ aN,
                         const double* __restrict__ aS, const double* __restrict__ nine, size_t n0,
                         const double* __restrict__ bv,
                         const unsigned char* __restrict__ act, const double* __restrict__ x,
                         double* __restrict__ xn, double* __restrict__ r, const double* __restrict__ r0,
                         const double* __restrict__ ph, const double* __restrict__ sh,
                         const double* __restrict__ sv, const double* __restrict__ t,
                         double* __restrict__ sc, double* __restrict__ part, int nb,
                         const double* __restrict__ p2, int n2, double* __restrict__ host)
{
    __shared__ double lds[3 * (BX * BY / 64)];
    CELL_IJ;
    const bool in = i <= ni && j <= nj;
    const size_t k = in ? (size_t)i * ld + j : (size_t)ld + 1;  // threads past the level read cell (1, 1)
    // this cell's values and, on the 5-point operator, the neighbours' x,
    // ph, sh and activity: loaded before om is known (om only scales them)
    const bool on = in && act[k] != 0;
    const double xk0 = x[k], phk = ph[k], shk = sh[k], svk = sv[k], tk = t[k], r0k = r0[k], bk = bv[k], pk = aP[k];
    const size_t nbr[4] = {k + ld, k - ld, k + 1, k - 1};
    double cpl[4] = {0.0, 0.0, 0.0, 0.0}, nx[4] = {0.0, 0.0, 0.0, 0.0}, nph[4] = {0.0, 0.0, 0.0, 0.0},
           nsh[4] = {0.0, 0.0, 0.0, 0.0};
    bool nact[4] = {false, false, false, false};
    if (!nine) {
        cpl[0] = aE[k];
        cpl[1] = aW[k];
        cpl[2] = aN[k];
        cpl[3] = aS[k];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            nact[q] = act[nbr[q]] != 0;
            nx[q] = x[nbr[q]];
            nph[q] = ph[nbr[q]];
            nsh[q] = sh[nbr[q]];
        }
    }
    double om;
    if (p2) {
        // om = (t, s) / (t, t) from the up kernel's tile partials, summed by
        // every workgroup in the same order; workgroup 0 stores it
        double tv[2] = {0.0, 0.0};
        for (int q = lin_tid(); q < n2; q += BX * BY) {
            tv[0] += p2[q];
            tv[1] += p2[n2 + q];
        }
        block_sum_vec_bcast<BX * BY, 2>(tv, lds);
        const double ts = tv[0], tt = tv[1];
        om = (tt > 0.0 && isfinite(ts / tt)) ? ts / tt : 0.0;
        if (lin_tid() == 0 && blockIdx.x == 0 && blockIdx.y == 0) {
            sc[SC_TS] = ts;
            sc[SC_TT] = tt;
            sc[SC_OM] = om;
        }
    } else {
        om = sc[SC_OM];
    }
    double sums[3] = {0.0, 0.0, 0.0};  // (r, r), (r0, r), |b - A x|^2
    if (in) {
        const double al = sc[SC_AL];
        double xk = 0.0;
        if (on) {
            xk = xk0 + al * phk + om * shk;
            const double q = svk - om * tk;
            r[k] = q;
            sums[0] = q * q;
            sums[1] = r0k * q;
            double off;
            if (!nine) {
                auto xnb = [&](int m) { return nact[m] ? nx[m] + al * nph[m] + om * nsh[m] : 0.0; };
                off = cpl[0] * xnb(0) + cpl[1] * xnb(1) + cpl[2] * xnb(2) + cpl[3] * xnb(3);
            } else {
                auto xnew = [&](size_t m) { return act[m] ? x[m] + al * ph[m] + om * sh[m] : 0.0; };
                off = offdiag(aE, aW, aN, aS, nine, n0, k, ld, xnew);
            }
            const double rt = bk - pk * xk + off;
            sums[2] = rt * rt;
        }
        xn[k] = xk;
    }
    block_sum_vec_bcast<BX * BY, 3>(sums, lds);
    if (lin_tid() == 0) {
        const double t1 = sums[0], t2 = sums[1], t3 = sums[2];
        const int bid = blockIdx.y * gridDim.x + blockIdx.x;
        part[bid] = t1;
        part[nb + bid] = t2;
        part[2 * nb + bid] = t3;
        if (host) {
            // the host's copy for the convergence test (pinned memory, read
            // after the stream is synchronised): the partials after the
            // scalar slots, and the two scalars it reads
            host[SC_COUNT + bid] = t1;
            host[SC_COUNT + nb + bid] = t2;
            host[SC_COUNT + 2 * nb + bid] = t3;
            if (bid == 0) {
                host[SC_BB] = sc[SC_BB];
                host[SC_SIDE] = sc[SC_SIDE];
                host[SC_D0] = sc[SC_D0];
                host[SC_DC] = sc[SC_DC];
            }
        }
    }
}

// finishing reductions (one block, fixed order)
__device__ inline double sum_part(const double* part, int nb, double* lds)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) s += part[k];
    return block_sum<256>(s, lds);
}

// what: 0 |b|^2 (+ mixed input scale); 1 after restart: rr, rho1 = (r0,r);
// 2 al = rho1 / (r0, v); 3 om = (t,s)/(t,t); 4 after the x update: rr, rho = rho1, rho1 = (r0,r)
__global__ void k_bicg_finish(const double* __restrict__ part, int nb, double* __restrict__ sc, int what,
                              double nact)
{
    __shared__ double lds[8];
    const double a = sum_part(part, nb, lds);
    const double b = (what == 1 || what == 3 || what == 4) ? sum_part(part + nb, nb, lds) : 0.0;
    const double c = what == 4 ? sum_part(part + 2 * nb, nb, lds) : 0.0;
    if (threadIdx.x != 0) return;
    switch (what) {
    case 0:
        sc[SC_BB] = a;
        sc[SC_SCALE] = a > 0.0 ? sqrt(nact / a) : 1.0;
        break;
    case 1:
        sc[SC_RR] = a;
        sc[SC_RT] = a;
        sc[SC_RHO1] = b;
        sc[SC_RHO] = 1.0;
        sc[SC_AL] = 1.0;
        sc[SC_OM] = 1.0;
        break;
    case 2: {
        sc[SC_R0V] = a;
        const double al = sc[SC_RHO1] / a;
        sc[SC_AL] = (a != 0.0 && isfinite(al)) ? al : 0.0;  // breakdown: no step
        break;
    }
    case 3:
        sc[SC_TS] = a;
        sc[SC_TT] = b;
        sc[SC_OM] = (b > 0.0 && isfinite(a / b)) ? a / b : 0.0;
        break;
    case 4:
        sc[SC_RR] = a;
        sc[SC_RHO] = sc[SC_RHO1];
        sc[SC_RHO1] = b;
        sc[SC_RT] = c;
        break;
    }
}

// ------------------------------------------------------------------ host side
MGSolver::~MGSolver()
{
    drop_graphs();
    for (auto& e : prof.ev) (void)hipEventDestroy(e);
    if (h_scal && h_owned) (void)hipHostFree(h_scal);
}

static int pad_ld(int nj) { return ((nj + 2 + 7) / 8) * 8; }

void MGSolver::configure(int ni, int nj, hipStream_t s)
{
    (void)s;
    if (ni == cfg_ni && nj == cfg_nj && !arena) return;
    arena = nullptr;
    arena_base = nullptr;
    build(ni, nj);
    const size_t nsc = scal.n;
    if (h_scal_n < nsc || !h_owned) {
        if (h_scal && h_owned) CATHODE_HIP(hipHostFree(h_scal));
        h_scal = nullptr;
        CATHODE_HIP(hipHostMalloc(&h_scal, nsc * sizeof(double), hipHostMallocDefault));
        h_scal_n = nsc;
        h_owned = true;
    }
    // plain fills on the null stream, which the library's non-blocking
    // streams do not wait for
    zero_buffers(nullptr);
    CATHODE_HIP(hipDeviceSynchronize());
}

size_t MGSolver::configure_in(int ni, int nj, unsigned char* base, size_t cap, double* hbuf, size_t hn,
                              hipStream_t s)
{
    size_t need, nsc;
    {
        // the bytes of this shape, counted by a solver that touches no memory
        MGSolver probe;
        Arena count;
        probe.arena = &count;
        probe.build(ni, nj);
        need = count.used;
        nsc = probe.scal.n;
    }
    if (need > cap) return need;
    if (nsc > hn) return SIZE_MAX;
    if (arena && arena_base == base && ni == cfg_ni && nj == cfg_nj) return need;  // unchanged
    if (h_scal && h_owned) CATHODE_HIP(hipHostFree(h_scal));
    region = Arena{base, 0};
    arena = &region;
    arena_base = base;
    h_scal = hbuf;
    h_scal_n = hn;
    h_owned = false;
    autotune_on = false;
    build(ni, nj);
    zero_buffers(s);
    return need;
}

void MGSolver::pull_host(size_t n, hipStream_t s)
{
    if (!arena) return;
    CATHODE_HIP(hipMemcpyAsync(h_scal, dmirror.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    CATHODE_HIP(hipStreamSynchronize(s));
}

void MGSolver::zero_buffers(hipStream_t s)
{
    const size_t tot = level_off[nl], n0 = level_off[1] - level_off[0];
    CATHODE_HIP(hipMemsetAsync(store64.p, 0, 8 * tot * sizeof(double), s));
    CATHODE_HIP(hipMemsetAsync(store32.p, 0, 8 * tot * sizeof(float), s));
    CATHODE_HIP(hipMemsetAsync(mass.p, 0, tot * sizeof(double), s));
    CATHODE_HIP(hipMemsetAsync(actall.p, 0, tot, s));
    CATHODE_HIP(hipMemsetAsync(act0.p, 0, n0, s));
    CATHODE_HIP(hipMemsetAsync(xalt.p, 0, n0 * sizeof(double), s));
    CATHODE_HIP(hipMemsetAsync(bsol.p, 0, n0 * sizeof(double), s));
    CATHODE_HIP(hipMemsetAsync(xsol.p, 0, n0 * sizeof(double), s));
    CATHODE_HIP(hipMemsetAsync(scal.p, 0, scal.n * sizeof(double), s));
    CATHODE_HIP(hipMemsetAsync(kv.p, 0, 8 * n0 * sizeof(double), s));
    if (arena) CATHODE_HIP(hipMemsetAsync(dmirror.p, 0, dmirror.n * sizeof(double), s));
}

void MGSolver::build(int ni, int nj)
{
    drop_graphs();  // they hold the old buffers
    setup_mode = setup_ni = setup_nj = -1;  // no hierarchy to reuse
    cfg_ni = ni;
    cfg_nj = nj;
    for (int q = 0; q < 2; q++) {
        tuned[q] = false;
        for (int l = 0; l < MG_MAX_LEVELS; l++) tiles[q][l] = 0;
    }
    // shapes
    int dims[MG_MAX_LEVELS][2];
    int n = 0;
    dims[0][0] = ni;
    dims[0][1] = nj;
    n = 1;
    while (n < MG_MAX_LEVELS) {
        int a = dims[n - 1][0], b = dims[n - 1][1];
        if ((long)a * b <= MG_COARSEST_MAX || (a == 1 && b == 1)) break;
        dims[n][0] = a > 1 ? (a + 1) / 2 : a;
        dims[n][1] = b > 1 ? (b + 1) / 2 : b;
        n++;
    }
    nl = n;
    size_t tot = 0;
    for (int l = 0; l < nl; l++) {
        level_off[l] = tot;
        tot += (size_t)(dims[l][0] + 2) * pad_ld(dims[l][1]);
    }
    level_off[nl] = tot;
    grab(store64, 8 * tot);
    grab(store32, 8 * tot);
    grab(mass, tot);
    grab(actall, tot);
    for (int l = 0; l < nl; l++) {
        const size_t o = level_off[l], sz = level_off[l + 1] - level_off[l];
        Level<double>& A = L64[l];
        Level<float>& B = L32[l];
        A.ni = B.ni = dims[l][0];
        A.nj = B.nj = dims[l][1];
        A.ld = B.ld = pad_ld(dims[l][1]);
        A.cx = B.cx = (l + 1 < nl) && dims[l][0] > 1;
        A.cy = B.cy = (l + 1 < nl) && dims[l][1] > 1;
        double* p = store64.p;
        float* q = store32.p;
        // b, x and xu of level l; the five coefficient arrays are placed below
        A.b = p + 5 * tot + o;
        A.x = p + 6 * tot + o;
        A.xu = p + 7 * tot + o;
        B.b = q + 5 * tot + o;
        B.x = q + 6 * tot + o;
        B.xu = q + 7 * tot + o;
        (void)sz;
    }
    // the coefficient block of level l starts at 5*level_off[l], so that the
    // five arrays of a level are contiguous (one FP32 conversion launch)
    for (int l = 0; l < nl; l++) {
        const size_t o = 5 * level_off[l], sz = level_off[l + 1] - level_off[l];
        L64[l].aP = store64.p + o;
        L64[l].aE = store64.p + o + sz;
        L64[l].aW = store64.p + o + 2 * sz;
        L64[l].aN = store64.p + o + 3 * sz;
        L64[l].aS = store64.p + o + 4 * sz;
        L32[l].aP = store32.p + o;
        L32[l].aE = store32.p + o + sz;
        L32[l].aW = store32.p + o + 2 * sz;
        L32[l].aN = store32.p + o + 3 * sz;
        L32[l].aS = store32.p + o + 4 * sz;
    }
    const size_t n0 = level_off[1] - level_off[0];
    grab(act0, n0);
    grab(bsol, n0);
    grab(xsol, n0);
    grab(xalt, n0);
    grab(w0, nj + 2);
    // pairs of fine levels in one launch each way (CATHODE_PAIR=0 turns this off)
    {
        const char* e = std::getenv("CATHODE_PAIR");
        pair_on = !(e && e[0] == '0');
        const char* c = std::getenv("CATHODE_PAIR_TILE");
        pair_cfg = (c && *c) ? std::atoi(c) : 3;  // 8 x 8 tiles of the second level, 1024 threads
    }
    // levels from lt on run inside the one-workgroup tail kernel
    lt = nl - 1;
    // the tail's first level may hold TAIL_TOP_MAX cells, every further one TAILR_NT
    auto cells = [&](int l) { return (long)dims[l][0] * dims[l][1]; };
    while (lt > 0 && cells(lt - 1) <= TAIL_TOP_MAX && cells(lt) <= TAILR_NT && nl - (lt - 1) <= TAIL_REG_LEVELS) lt--;
    tail_cells = 0;
    for (int l = lt; l < nl; l++) tail_cells += (dims[l][0] + 2) * (dims[l][1] + 2);
    {
        int tiles_max = 0;
        for (int cfg = 0; cfg < VT_NCFG; cfg++) tiles_max = std::max(tiles_max, vtile_count(cfg, ni, nj));
        grab(part, std::max(3 * (cdiv(nj, BX) * cdiv(ni, BY)), 2 * tiles_max) + 64);
    }
    // the scalar slots, then three partial vectors of the 2-D block grid
    const size_t nsc = SC_COUNT + 3 * (size_t)(cdiv(nj, BX) * cdiv(ni, BY)) + 64;
    grab(scal, nsc);
    grab(kv, 9 * n0);
    grab(cinv64, MG_COARSEST_MAX * MG_COARSEST_MAX);
    grab(presum, PS_COUNT + 2);
    if (arena) grab(dmirror, nsc);
    // V-cycle cut candidates: level 0 and the last level of the first fine
    // launch unit below it (levels 1 and 2, which run as a pair unless
    // CATHODE_PAIR=0, or level 1 alone), which the setup launch builds; the
    // same levels with pairs on or off, so both give the same results
    tcut = -1;
    if (lt > 1) tcut = (2 < lt && L64[1].cx && L64[1].cy) ? 2 : 1;
    dpart_n = (cdiv(ni, 16) + 2) * (cdiv(nj, 32) + 2);  // bounds the setup launch's workgroups
    grab(dpart, 2 * (size_t)dpart_n);
}

CoarsenArgs MGSolver::coarsen_args(int l, const double* w)
{
    Level<double>& A = L64[l];
    Level<double>& C = L64[l + 1];
    Level<float>& G = L32[l + 1];
    CoarsenArgs a;
    a.fni = A.ni; a.fnj = A.nj; a.fld = A.ld; a.fi = A.cx ? 2 : 1; a.fj = A.cy ? 2 : 1;
    a.faP = A.aP; a.faE = A.aE; a.faW = A.aW; a.faN = A.aN; a.faS = A.aS;
    a.fm = mass.p + level_off[l];
    a.fact = actall.p + level_off[l];
    a.w = l == 0 ? w : nullptr;
    a.cni = C.ni; a.cnj = C.nj; a.cld = C.ld;
    a.caP = C.aP; a.caE = C.aE; a.caW = C.aW; a.caN = C.aN; a.caS = C.aS; a.cb = C.b; a.cx = C.x;
    a.cm = mass.p + level_off[l + 1];
    a.cact = actall.p + level_off[l + 1];
    a.gP = G.aP; a.gE = G.aE; a.gW = G.aW; a.gN = G.aN; a.gS = G.aS;
    a.hx = A.cx ? 0.5 : 1.0;
    a.hy = A.cy ? 0.5 : 1.0;
    a.hb = (A.cx && A.cy) ? 0.5 : ((A.cx || A.cy) ? 0.75 : 1.0);
    return a;
}

// Builds the hierarchy for the level-0 system the caller filled in, plus the
// setup scalars (|b|^2, FP32 scales): 3 launches while lt <= 3.
void MGSolver::setup(int mode, hipStream_t s, const WarmSet& ws)
{
    Level<double>& F = L64[0];
    const size_t n0 = level_off[1] - level_off[0];
    const int mixed = mode == CATHODE_MODE_MIXED ? 1 : 0;
    const int nb = cdiv(F.nj, BX) * cdiv(F.ni, BY);
    const bool act_changed = act_dirty;
    // level-0 activity lives in act0 (filled by the caller); mirror into actall
    if (act_dirty) {
        CATHODE_HIP(hipMemcpyAsync(actall.p, act0.p, n0, hipMemcpyDeviceToDevice, s));
        act_dirty = false;
    }
    // the caller's promise that only b changed since the last setup, for the
    // same precision: the levels, FP32 copies and coarsest inverse stay
    const bool rhs_only = same_operator && setup_mode == mode && !act_changed && cfg_ni == setup_ni &&
                          cfg_nj == setup_nj;
    same_operator = false;
    setup_mode = mode;
    setup_ni = cfg_ni;
    setup_nj = cfg_nj;
    double* sbuf = F.b;  // row excess of level 0 (F.b is rewritten by the outer loop)
    if (!prepared) {
        hipLaunchKernelGGL(k_l0_prepare, grid2(F.ni, F.nj), dim3(BX, BY), 0, s, F.ni, F.nj, F.ld, act0.p, F.aP, F.aE,
                           F.aW, F.aN, F.aS, bsol.p, sbuf, part.p, nb);
        CATHODE_LAUNCH_CHECK();
    }
    prepared = false;
    const double* w = user_w;
    if (!w && !rhs_only) {
        hipLaunchKernelGGL(k_l0_weights, dim3(1), dim3(256), (F.nj + 2) * sizeof(double), s, F.ni, F.nj, F.ld, F.aN,
                           F.aS, w0.p);
        CATHODE_LAUNCH_CHECK();
    }
    if (!w) w = w0.p;
    // scalars, level-0 mass and the first coarsenings: one launch; a
    // workgroup owns a 16 x 32 block of level 0 (fewer rows / columns where a
    // direction stops coarsening)
    const int nf = std::min(std::min(lt, nl - 1), SETUP_FUSE_MAX);
    {
        Level<float>& G = L32[0];
        SetupLevels a{};
        a.ni = F.ni; a.nj = F.nj; a.ld = F.ld; a.act = act0.p; a.s = sbuf; a.w = w; a.m = mass.p + level_off[0];
        a.aP = F.aP; a.aE = F.aE; a.aW = F.aW; a.aN = F.aN; a.aS = F.aS;
        a.fP = G.aP; a.fE = G.aE; a.fW = G.aW; a.fN = G.aN; a.fS = G.aS;
        a.part = presummed(part.p, nb, 2, 1u, PS_SETUP, a.nb, s);  // max|aP| and |b|^2 partials
        a.nact = (double)F.ni * (double)F.nj;
        a.nf = nf;
        a.pi[nf] = a.pj[nf] = 1;
        for (int l = nf - 1; l >= 0; l--) {
            a.pi[l] = a.pi[l + 1] * (L64[l].cx ? 2 : 1);
            a.pj[l] = a.pj[l + 1] * (L64[l].cy ? 2 : 1);
            a.L[l] = coarsen_args(l, w);
        }
        a.ri = std::max(1, 16 / a.pi[0]);
        a.rj = std::max(1, 32 / a.pj[0]);
        const int cni = L64[nf].ni, cnj = L64[nf].nj;
        const dim3 sg(cdiv(cnj, a.rj), cdiv(cni, a.ri));
        a.ws = ws;
        a.b = bsol.p;
        if (ws.k > 0) {
            warm_nb = (int)(sg.x * sg.y);
            wpart.alloc((size_t)WARM_NP * warm_nb + 64);
            a.wpart = wpart.p;
        }
        a.rhs_only = rhs_only ? 1 : 0;
        // the cut candidates' excess fractions (reduced by the first restart)
        setup_dnb = (int)(sg.x * sg.y);
        const bool dmeas = !rhs_only && setup_dnb <= dpart_n;
        a.dpart = dmeas ? dpart.p : nullptr;
        a.dnb = dpart_n;
        a.dcl = (tcut >= 1 && tcut <= nf) ? tcut : -1;
        if (!rhs_only) {
            dpart_ok = dmeas;
            setup_dcl = a.dcl;
        }
        hipLaunchKernelGGL(k_setup_levels, sg, dim3(SETUP_NT), 0, s, a, scal.p, mixed);
        CATHODE_LAUNCH_CHECK();
    }
    if (rhs_only) return;  // the levels and the coarsest inverse of the last setup stand
    // further fine levels: one launch each; the rest in one workgroup
    for (int l = nf; l + 1 < nl && l < lt; l++) {
        Level<double>& C = L64[l + 1];
        hipLaunchKernelGGL(k_coarsen, grid2(C.ni, C.nj), dim3(BX, BY), 0, s, coarsen_args(l, w), (const double*)scal.p,
                           mixed);
    }
    CoarseInv ci{};
    {
        const Level<double>& Z = L64[nl - 1];
        ci.ni = Z.ni; ci.nj = Z.nj; ci.ld = Z.ld;
        ci.aP = Z.aP; ci.aE = Z.aE; ci.aW = Z.aW; ci.aN = Z.aN; ci.aS = Z.aS;
        ci.act = actall.p + level_off[nl - 1];
        ci.inv64 = cinv64.p;
    }
    bool inverted = false;
    for (int l0 = lt; l0 + 1 < nl; l0 += CHAIN_MAX) {
        CoarsenChain ch;
        ch.n = 0;
        for (int l = l0; l + 1 < nl && ch.n < CHAIN_MAX; l++) ch.L[ch.n++] = coarsen_args(l, w);
        const bool last = l0 + ch.n + 1 >= nl;
        CoarseInv cc = ci;
        if (!last) cc.inv64 = nullptr;
        if (last) {
            chain_pending = true;
            pend_mixed = mixed;
            pend_ch = ch;
            pend_ci = cc;
        } else {
            hipLaunchKernelGGL(k_coarsen_chain, dim3(1), dim3(512), 0, s, ch, (const double*)scal.p, mixed, cc);
        }
        inverted = inverted || last;
    }
    if (!inverted) hipLaunchKernelGGL(k_coarse_inverse, dim3(1), dim3(CINV_NT), 0, s, ci, (const double*)scal.p);
    CATHODE_LAUNCH_CHECK();
}

// cut >= 0: the V-cycle stops at fine level cut (the last level of a launch
// unit): that level is smoothed from x = 0 by its down and up launches with
// no coarse correction, and the coarser levels and the tail are skipped.
template <class T>
void MGSolver::vcycle(Level<T>* L, hipStream_t s, const VFuse& fdown, const VFuse& fup, int cut)
{
    const int last = cut >= 0 ? cut : lt - 1;  // the last fine level visited
    const double* w = user_w ? user_w : w0.p;
    auto vlevel = [&](int l) {
        Level<T>& A = L[l];
        return VLevel<T>{A.ni, A.nj, A.ld, A.aP, A.aE, A.aW, A.aN, A.aS, actall.p + level_off[l], A.b, A.x, A.xu};
    };
    // fine levels: fused down kernels; below the finest, a level that
    // coarsens 2x2 into a next fine level runs with it (pair_start[l])
    bool pair_start[MG_MAX_LEVELS] = {};
    for (int l = 0; l <= last;) {
        Level<T>& A = L[l];
        if (pair_on && l >= 1 && l + 1 < lt && A.cx && A.cy) {
            Level<T>& C = L[l + 1];
            Level<T>& D = L[l + 2];
            pair_start[l] = true;
            launch_vpair<T>(pair_cfg, true, s, vlevel(l), vlevel(l + 1), C.cx ? 2 : 1, C.cy ? 2 : 1, D.ni, D.nj, D.ld,
                            D.b);
            l += 2;
            continue;
        }
        Level<T>& C = L[l + 1];
        VLevel<T> V = vlevel(l);
        V.xo = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (prof.enabled && l == 0) prof_events(e0, e1);
        launch_vtile<T, true>(tile_of<T>(l), l == 0, s, V, A.cx ? 2 : 1, A.cy ? 2 : 1, w, C.ni, C.nj, C.ld, C.b,
                              l == 0 ? fdown : VFuse(), e0, e1);
        l++;
    }
    // coarse levels in one workgroup
    if (cut < 0) {
        TailArgs<T> ta{};
        ta.nt = nl - lt;
        int off = 0;
        for (int q = 0; q < ta.nt; q++) {
            Level<T>& A = L[lt + q];
            ta.ni[q] = A.ni;
            ta.nj[q] = A.nj;
            ta.off[q] = off;
            off += (A.ni + 2) * (A.nj + 2);
            ta.cx[q] = A.cx;
            ta.cy[q] = A.cy;
            ta.gld[q] = A.ld;
            ta.aP[q] = A.aP;
            ta.aE[q] = A.aE;
            ta.aW[q] = A.aW;
            ta.aN[q] = A.aN;
            ta.aS[q] = A.aS;
            ta.act[q] = actall.p + level_off[lt + q];
        }
        ta.ncell = off;
        ta.b_in = L[lt].b;
        ta.x_out = L[lt].x;
        ta.w = lt == 0 ? w : nullptr;
        ta.cn = L[nl - 1].ni * L[nl - 1].nj;
        // the coarsest level is solved in FP64 in both precisions: the FP32
        // levels hold s32 A, so their solution is (A^-1 b) / s32
        ta.cinv = cinv64.p;
        ta.s32 = sizeof(T) == 4 ? scal.p + SC_S32 : nullptr;
        const int top = L[lt].ni * L[lt].nj;
        const size_t lds = tail_lds_bytes<T>(off, top);
        if (top <= TAILR_NT) launch_tail<T, 1>(ta, lds, s);
        else launch_tail<T, 2>(ta, lds, s);
    }
    // fine levels: fused up kernels; the coarse correction is the next
    // level's final x (xu of a fine level or of a pair's first level, x of
    // the tail's first level)
    for (int l = last; l >= 0;) {
        if (l >= 1 && pair_start[l - 1]) {
            Level<T>& C = L[l];
            Level<T>& D = L[l + 1];
            launch_vpair<T>(pair_cfg, false, s, vlevel(l - 1), vlevel(l), C.cx ? 2 : 1, C.cy ? 2 : 1, D.ni, D.nj, D.ld,
                            l == cut ? nullptr : (l + 1 < lt ? D.xu : D.x));
            l -= 2;
            continue;
        }
        Level<T>& A = L[l];
        Level<T>& C = L[l + 1];
        launch_vtile<T, false>(tile_of<T>(l), l == 0, s, vlevel(l), A.cx ? 2 : 1, A.cy ? 2 : 1, (const double*)nullptr,
                               C.ni, C.nj, C.ld, l == cut ? nullptr : (l + 1 < lt ? C.xu : C.x),
                               l == 0 ? fup : VFuse());
        l--;
    }
    CATHODE_LAUNCH_CHECK();
}

void MGSolver::flush_chain(hipStream_t s)
{
    if (!chain_pending) return;
    chain_pending = false;
    hipLaunchKernelGGL(k_coarsen_chain, dim3(1), dim3(512), 0, s, pend_ch, (const double*)scal.p, pend_mixed, pend_ci);
    CATHODE_LAUNCH_CHECK();
}

void MGSolver::prof_events(hipEvent_t& a, hipEvent_t& b)
{
    CATHODE_HIP(hipEventCreate(&a));
    CATHODE_HIP(hipEventCreate(&b));
    prof.ev.push_back(a);
    prof.ev.push_back(b);
}

double MGSolver::read_scalar(const double* d, hipStream_t s)
{
    if (arena) CATHODE_HIP(hipMemcpyAsync(h_scal, d, sizeof(double), hipMemcpyDeviceToHost, s));
    else fetch_host(h_scal, d, 1, s);
    CATHODE_HIP(hipStreamSynchronize(s));
    return h_scal[0];
}

// the tail kernel with C0 first-level cells per thread; LDS beyond the
// default 64 KiB is enabled once per instantiation
template <class T, int C0>
static void launch_tail(const TailArgs<T>& ta, size_t lds, hipStream_t s)
{
    static bool big_lds = false;
    if (lds > 65536 && !big_lds) {
        CATHODE_HIP(hipFuncSetAttribute((const void*)k_tail_reg<T, C0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        160 * 1024));
        big_lds = true;
    }
    hipLaunchKernelGGL((k_tail_reg<T, C0>), dim3(1), dim3(TAILR_NT), lds, s, ta);
}

template <class T>
float MGSolver::time_tile(Level<T>* L, int l, int cfg, hipStream_t s)
{
    Level<T>& A = L[l];
    Level<T>& C = L[l + 1];
    const double* w = user_w ? user_w : w0.p;
    VLevel<T> V{A.ni, A.nj, A.ld, A.aP, A.aE, A.aW, A.aN, A.aS, actall.p + level_off[l], A.b, A.x, A.xu};
    hipEvent_t e0, e1;
    CATHODE_HIP(hipEventCreate(&e0));
    CATHODE_HIP(hipEventCreate(&e1));
    // one untimed pair (code load), then 4 timed pairs
    float ms = 0.f;
    for (int rep = 0; rep < 5; rep++) {
        if (rep == 1) CATHODE_HIP(hipEventRecord(e0, s));
        launch_vtile<T, true>(cfg, l == 0, s, V, A.cx ? 2 : 1, A.cy ? 2 : 1, w, C.ni, C.nj, C.ld, C.b);
        launch_vtile<T, false>(cfg, l == 0, s, V, A.cx ? 2 : 1, A.cy ? 2 : 1, (const double*)nullptr, C.ni, C.nj,
                               C.ld, C.x);
    }
    CATHODE_HIP(hipEventRecord(e1, s));
    CATHODE_LAUNCH_CHECK();
    CATHODE_HIP(hipEventSynchronize(e1));
    CATHODE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

// Tile choices per (shape, precision), shared by every solver of the process:
// a shape is timed once, later solvers of the same shape reuse the result.
struct TileCacheEntry {
    int ni, nj, prec;
    int tiles[MG_MAX_LEVELS];
};
static std::mutex g_tile_mu;
static std::vector<TileCacheEntry> g_tile_cache;

void MGSolver::autotune(int mixed, hipStream_t s)
{
    const int q = mixed ? 1 : 0;
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        for (const auto& e : g_tile_cache)
            if (e.ni == cfg_ni && e.nj == cfg_nj && e.prec == q) {
                for (int l = 0; l < MG_MAX_LEVELS; l++) tiles[q][l] = e.tiles[l];
                tuned[q] = true;
                return;
            }
    }
    for (int l = 0; l < lt; l++) {
        int best = 0;
        float tbest = 0.f;
        for (int cfg = 0; cfg < VT_NCFG; cfg++) {
            // the down and up kernels overwrite only V-cycle scratch (x of this
            // level, b of the next), which every V-cycle rewrites before use
            const float t = mixed ? time_tile<float>(L32, l, cfg, s) : time_tile<double>(L64, l, cfg, s);
            if (cfg == 0 || t < tbest) {
                best = cfg;
                tbest = t;
            }
        }
        tiles[q][l] = best;
        if (std::getenv("CATHODE_VERBOSE"))
            std::fprintf(stderr, "[cathode] autotune %s level %d (%dx%d): tile %dx%d/%d, %.2f us per down+up\n",
                         mixed ? "fp32" : "fp64", l, L64[l].ni, L64[l].nj, VT_CFGS[best].ti, VT_CFGS[best].tj,
                         VT_CFGS[best].nt, 1e3 * tbest / 4);
    }
    tuned[q] = true;
    TileCacheEntry e{cfg_ni, cfg_nj, q, {}};
    for (int l = 0; l < MG_MAX_LEVELS; l++) e.tiles[l] = tiles[q][l];
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile_cache.push_back(e);
}

void MGSolver::enable_nine(bool on)
{
    use_nine = on;
    if (on) nine.alloc(8 * n0());
}

void MGSolver::lump_nine(hipStream_t s)
{
    const Level<double>& F = L64[0];
    hipLaunchKernelGGL(k_lump9, grid2(F.ni, F.nj), dim3(BX, BY), 0, s, F.ni, F.nj, F.ld, (const double*)nine.p, n0(),
                       (const unsigned char*)act0.p, F.aE, F.aW, F.aN, F.aS);
    CATHODE_LAUNCH_CHECK();
}

void MGSolver::tile_shape(int mixed, int l, int* ti, int* tj, int* nt) const
{
    int a = 0, b = 0, c = 0;
    if (l >= 0 && l < lt) {
        const TileCfg& G = VT_CFGS[tiles[mixed ? 1 : 0][l]];
        a = G.ti;
        b = G.tj;
        c = G.nt;
    }
    if (ti) *ti = a;
    if (tj) *tj = b;
    if (nt) *nt = c;
}

void MGSolver::drop_graphs()
{
    for (auto& gr : graphs)
        if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
    graphs.clear();
}

const MGSolver::IterGraph* MGSolver::find_graph(int mixed, int parity, const double* xa, const double* pa,
                                                int cut) const
{
    const double* w = user_w ? user_w : w0.p;
    for (const auto& gr : graphs) {
        if (gr.mixed != mixed || gr.parity != parity || gr.xa != xa || gr.pa != pa || gr.w != w || gr.cut != cut)
            continue;
        if (std::memcmp(gr.tiles, tiles[mixed], sizeof(gr.tiles)) != 0) continue;
        return &gr;
    }
    return nullptr;
}

// Captures what enqueue launches on s into an executable graph.  A failed
// capture or instantiation leaves the stream usable and returns nullptr (the
// caller then launches eagerly); a failing launch inside is rethrown after
// the capture has been ended.
template <class Fn>
const MGSolver::IterGraph* MGSolver::capture_graph(int mixed, int parity, const double* xa, const double* pa,
                                                   int cut, hipStream_t s, Fn&& enqueue)
{
    if (graphs.size() >= 8) drop_graphs();
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    bool failed = false;
    hip_failure err{hipSuccess};
    try {
        enqueue();
    } catch (const hip_failure& f) {
        failed = true;
        err = f;
    }
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &gr);
    if (failed || e != hipSuccess || !gr) {
        if (gr) (void)hipGraphDestroy(gr);
        (void)hipGetLastError();
        if (failed) throw err;
        return nullptr;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ei != hipSuccess || !ex) {
        (void)hipGetLastError();
        return nullptr;
    }
    IterGraph ig;
    ig.exec = ex;
    ig.mixed = mixed;
    ig.parity = parity;
    ig.cut = cut;
    ig.xa = xa;
    ig.pa = pa;
    ig.w = user_w ? user_w : w0.p;
    std::memcpy(ig.tiles, tiles[mixed], sizeof(ig.tiles));
    graphs.push_back(ig);
    return &graphs.back();
}

int MGSolver::solve(int maxITR, double tol, int mode, hipStream_t s, cathode_status* st)
{
    auto t0 = std::chrono::steady_clock::now();
    Level<double>& F = L64[0];
    const size_t n0 = level_off[1] - level_off[0];
    const int nb = cdiv(F.nj, BX) * cdiv(F.ni, BY);
    const int mixed = mode == CATHODE_MODE_MIXED ? 1 : 0;
    const dim3 g = grid2(F.ni, F.nj), blk(BX, BY);
    // the caller's warm start (not with a kept guess or the nine-point
    // operator): its dot products are formed by the setup launch
    WarmSet ws = (keep_guess || use_nine) ? WarmSet() : warm;
    warm = WarmSet();
    setup(mode, s, ws);
    // the V-cycle cut of this solve (the caller's hint, when this hierarchy
    // offers it)
    const int cut = cut_valid(cut_hint) ? cut_hint : -1;
    cut_hint = -1;
    cut_nx = -1;
    bool pulled = false;
    {
        const int q = mixed ? 1 : 0;
        const char* force = std::getenv("CATHODE_TILE");
        if (force && *force) {
            const int cfg = std::atoi(force);
            for (int l = 0; l < MG_MAX_LEVELS; l++) tiles[q][l] = (cfg >= 0 && cfg < VT_NCFG) ? cfg : 0;
            tuned[q] = false;
        } else if (autotune_on && !tuned[q]) {
            autotune(mixed, s);
        }
    }
    // without a kept guess the first restart zeroes x (interior; ghosts stay
    // 0), or sets the caller's warm start
    bool zero_guess = !keep_guess && ws.k == 0;
    keep_guess = false;
    double* r = kv.p;
    double* r0 = r + n0;
    double* p = r0 + n0;
    double* v = p + n0;
    double* sv = v + n0;
    double* t = sv + n0;
    double* ph = t + n0;
    double* sh = ph + n0;
    double* pa = p;         // current search direction
    double* pb = sh + n0;   // the next one (fused path: never the buffer being read)
    // BiCGStab work fused into the finest V-cycle kernels: needs a finest
    // level handled by the tile kernels and the 5-point operator
    const bool fused = !use_nine && lt > 0;
    const int ntiles = vtile_count(mixed ? tiles[1][0] : tiles[0][0], F.ni, F.nj);
    VFuse fbase;
    fbase.mixed = mixed;
    fbase.s_s32 = SC_S32;
    fbase.s_scale = SC_SCALE;
    fbase.s_rho1 = SC_RHO1;
    fbase.s_rho = SC_RHO;
    fbase.s_al = SC_AL;
    fbase.s_om = SC_OM;
    fbase.r = r;
    fbase.v = v;
    fbase.sc = scal.p;
    fbase.aP = F.aP;
    fbase.aE = F.aE;
    fbase.aW = F.aW;
    fbase.aN = F.aN;
    fbase.aS = F.aS;
    fbase.part = part.p;
    aborted = false;
    const double* n9 = use_nine ? nine.p : nullptr;
    double* xa = xsol.p;   // current iterate
    double* xb = xalt.p;   // next iterate (double buffer)
    int it = 0;
    double rel = 0.0, bb = -1.0;
    bool restart = true;
    // partials of the restart and x-update kernels: right after the scalar
    // slots, so one copy brings both to the host
    double* part3 = scal.p + SC_COUNT;
    // One fused iteration: p and the V-cycle input in the finest down kernel
    // (rho1 from the x-update / restart partials); ph, v = A ph and (r0, v)
    // partials in the finest up kernel; s = r - al v (al from those
    // partials) and the second V-cycle; sh, t = A sh, (t, s), (t, t); then
    // the x update, which takes om from the last partials.  rho1 alternates
    // between two slots, so a down kernel reads the previous one while
    // workgroup 0 stores the new.  Everything it launches depends only on
    // (mode, iteration parity, x / p buffers), so it can be captured.
    auto fused_iter = [&](int first, double* xa_, double* xb_, double* pa_, double* pb_) {
        const int s_new = (it & 1) ? SC_RHO1 : SC_RHO, s_prev = (it & 1) ? SC_RHO : SC_RHO1;
        VFuse fd = fbase, fu = fbase;
        fd.pro = 1;
        fd.first = first;
        fd.pin = pa_;
        fd.ps = pb_;
        fd.pred = presummed(part3 + nb, nb, 1, 0u, PS_RHO1, fd.npred, s);  // (r0, r)
        fd.s_rho1 = s_new;
        fd.s_rho = s_prev;
        fu.epi = 1;
        fu.y = ph;
        fu.out = v;
        fu.dotvec = r0;
        fu.want_tt = 0;
        if (mixed) vcycle<float>(L32, s, fd, fu, cut);
        else vcycle<double>(L64, s, fd, fu, cut);
        fd.pro = 2;
        fd.first = 0;
        fd.pin = nullptr;
        fd.ps = sv;
        fd.pred = presummed(part.p, ntiles, 1, 0u, PS_R0V, fd.npred, s);  // (r0, v)
        fd.s_r0v = SC_R0V;
        fu.y = sh;
        fu.out = t;
        fu.dotvec = sv;
        fu.want_tt = 1;
        if (mixed) vcycle<float>(L32, s, fd, fu, cut);
        else vcycle<double>(L64, s, fd, fu, cut);
        int n2 = 0;
        const double* p2 = presummed(part.p, ntiles, 2, 0u, PS_OM, n2, s);  // (t, s), (t, t)
        hipLaunchKernelGGL(k_bicg_x, g, blk, 0, s, F.ni, F.nj, F.ld, F.aP, F.aE, F.aW, F.aN, F.aS, n9, n0, bsol.p,
                           act0.p, xa_, xb_, r, r0, ph, sh, sv, t, scal.p, part3, nb, p2, n2, host_dst());
        CATHODE_LAUNCH_CHECK();
    };
    // V-cycle output on level 0: the finest up kernel's xu (x when the
    // whole hierarchy runs in the tail)
    const float* vout32 = lt > 0 ? L32[0].xu : L32[0].x;
    const double* vout64 = lt > 0 ? F.xu : F.x;
    // graph replay of the later iterations: CATHODE_GRAPHS=1 (off by
    // default: on ROCm 7 one hipGraphLaunch of the ~14 nodes costs more than
    // launching them one by one)
    const int unchecked = std::min(expected_its - 1, maxITR - 1);
    expected_its = 0;
    bool checked = false;
    const char* genv = std::getenv("CATHODE_GRAPHS");
    const bool graphs_on = fused && !prof.enabled && !arena && genv && genv[0] == '1';
    while (it < maxITR) {
        if (restart) {
            RestartArgs ra{};
            ra.ni = F.ni; ra.nj = F.nj; ra.ld = F.ld;
            ra.aP = F.aP; ra.aE = F.aE; ra.aW = F.aW; ra.aN = F.aN; ra.aS = F.aS; ra.nine = n9; ra.n0 = n0;
            ra.b = bsol.p; ra.x = xa; ra.act = act0.p;
            ra.r = r; ra.r0 = r0; ra.p = pa; ra.v = v; ra.part = part3; ra.nb = nb;
            ra.xout = (zero_guess || ws.k > 0) ? xa : nullptr;
            ra.ws = ws;
            if (ws.k > 0) ra.wpart = presummed(wpart.p, warm_nb, WARM_NP, 0u, PS_WARM, ra.wnb, s);
            ra.dpart = dpart_ok ? dpart.p : nullptr; ra.dnb = setup_dnb; ra.dstride = dpart_n; ra.sc = scal.p;
            // a cut V-cycle reads no tail level: the chain stays pending
            // until a solve on this hierarchy needs it
            if (chain_pending && cut < 0) {
                chain_pending = false;
                hipLaunchKernelGGL(k_restart_chain, dim3(1 + cdiv(g.x * g.y, 2)), dim3(BX, RC_NT / BX), 0, s, ra,
                                   pend_ch, (const double*)scal.p, pend_mixed, pend_ci);
            } else {
                hipLaunchKernelGGL(k_bicg_restart, g, blk, 0, s, ra);
            }
            zero_guess = false;
            ws = WarmSet();
            // the fused path sums these partials in its first down kernel
            if (!fused) hipLaunchKernelGGL(k_bicg_finish, dim3(1), dim3(256), 0, s, part3, nb, scal.p, 1, 0.0);
        }
        it++;
        if (fused) {
            const int first = restart ? 1 : 0;
            if (!first && graphs_on) {
                // later iterations replay a captured copy of the same launches
                const IterGraph* gr = find_graph(mixed, it & 1, xa, pa, cut);
                if (!gr) gr = capture_graph(mixed, it & 1, xa, pa, cut, s, [&] { fused_iter(first, xa, xb, pa, pb); });
                if (gr) CATHODE_HIP(hipGraphLaunch(gr->exec, s));
                else fused_iter(first, xa, xb, pa, pb);
            } else {
                fused_iter(first, xa, xb, pa, pb);
            }
            std::swap(pa, pb);
            restart = false;
        } else {
        // p, ph = M^-1 p, v = A ph, al
        if (mixed) {
            hipLaunchKernelGGL(k_bicg_p<float>, g, blk, 0, s, F.ni, F.nj, F.ld, act0.p, r, v, pa, L32[0].b,
                               L32[0].x, scal.p, restart ? 1 : 0, 1);
            vcycle<float>(L32, s);
            hipLaunchKernelGGL(k_bicg_apply<float>, g, blk, 0, s, F.ni, F.nj, F.ld, F.aP, F.aE, F.aW, F.aN,
                               F.aS, n9, n0, act0.p, vout32, ph, v, r0, scal.p, 1, 0, part.p, nb);
        } else {
            hipLaunchKernelGGL(k_bicg_p<double>, g, blk, 0, s, F.ni, F.nj, F.ld, act0.p, r, v, pa, F.b, F.x,
                               scal.p, restart ? 1 : 0, 0);
            vcycle<double>(L64, s);
            hipLaunchKernelGGL(k_bicg_apply<double>, g, blk, 0, s, F.ni, F.nj, F.ld, F.aP, F.aE, F.aW, F.aN,
                               F.aS, n9, n0, act0.p, vout64, ph, v, r0, scal.p, 0, 0, part.p, nb);
        }
        hipLaunchKernelGGL(k_bicg_finish, dim3(1), dim3(256), 0, s, part.p, nb, scal.p, 2, 0.0);
        restart = false;
        // s, sh = M^-1 s, t = A sh, om
        if (mixed) {
            hipLaunchKernelGGL(k_bicg_s<float>, g, blk, 0, s, F.ni, F.nj, F.ld, act0.p, r, v, sv, L32[0].b,
                               L32[0].x, scal.p, 1);
            vcycle<float>(L32, s);
            hipLaunchKernelGGL(k_bicg_apply<float>, g, blk, 0, s, F.ni, F.nj, F.ld, F.aP, F.aE, F.aW, F.aN,
                               F.aS, n9, n0, act0.p, vout32, sh, t, sv, scal.p, 1, 1, part.p, nb);
        } else {
            hipLaunchKernelGGL(k_bicg_s<double>, g, blk, 0, s, F.ni, F.nj, F.ld, act0.p, r, v, sv, F.b, F.x,
                               scal.p, 0);
            vcycle<double>(L64, s);
            hipLaunchKernelGGL(k_bicg_apply<double>, g, blk, 0, s, F.ni, F.nj, F.ld, F.aP, F.aE, F.aW, F.aN,
                               F.aS, n9, n0, act0.p, vout64, sh, t, sv, scal.p, 0, 1, part.p, nb);
        }
        hipLaunchKernelGGL(k_bicg_finish, dim3(1), dim3(256), 0, s, part.p, nb, scal.p, 3, 0.0);
        // x, recursive and true residuals
        hipLaunchKernelGGL(k_bicg_x, g, blk, 0, s, F.ni, F.nj, F.ld, F.aP, F.aE, F.aW, F.aN, F.aS, n9, n0, bsol.p,
                           act0.p, xa, xb, r, r0, ph, sh, sv, t, scal.p, part3, nb, (const double*)nullptr, 0,
                           host_dst());
        hipLaunchKernelGGL(k_bicg_finish, dim3(1), dim3(256), 0, s, part3, nb, scal.p, 4, 0.0);
        }
        CATHODE_LAUNCH_CHECK();
        std::swap(xa, xb);
        // the caller's expected iteration count: the iterations before the
        // last expected one run without a check (no round trip)
        if (it < unchecked) continue;
        // one host round trip per iteration: the x update stored its partials
        // and |b|^2 and the side slot in h_scal; |r|^2 and |b - A x|^2 are
        // summed here in a fixed order
        CATHODE_HIP(hipStreamSynchronize(s));
        pull_host(SC_COUNT + 3 * (size_t)nb, s);
        pulled = true;
        {
            double rr = 0.0, rt2 = 0.0;
            const double* hp = h_scal + SC_COUNT;
            for (int q = 0; q < nb; q++) {
                rr += hp[q];
                rt2 += hp[2 * nb + q];
            }
            h_scal[SC_RR] = rr;
            h_scal[SC_RT] = rt2;
        }
        bb = h_scal[SC_BB];
        const bool first_check = !checked;
        checked = true;
        if (first_check && h_scal[SC_SIDE] < side_cnv) {
            // the caller's stopping test (read with this round trip) says the
            // solve is not wanted: x is left as it is, the caller rolls back
            aborted = true;
            break;
        }
        if (!(bb > 0.0)) {
            rel = bb == 0.0 ? 0.0 : NAN;  // b = 0: HYPRE returns x = 0 (the step above added 0)
            if (bb == 0.0) it = 0;
            break;
        }
        const double rt = sqrt(h_scal[SC_RT] / bb);
        rel = rt;
        if (!std::isfinite(rel)) break;
        if (rt <= tol) break;
        // the recursion has drifted from the true residual: restart from it
        if (sqrt(h_scal[SC_RR] / bb) <= tol) restart = true;
    }
    if (maxITR <= 0) {
        // no iteration: x = 0, relative residual 1 (0 when b = 0)
        if (zero_guess || ws.k > 0) CATHODE_HIP(hipMemsetAsync(xsol.p, 0, n0 * sizeof(double), s));
        if (arena) CATHODE_HIP(hipMemcpyAsync(h_scal, scal.p, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
        else fetch_host(h_scal, scal.p, SC_COUNT, s);
        CATHODE_HIP(hipStreamSynchronize(s));
        rel = h_scal[SC_BB] > 0.0 ? 1.0 : 0.0;
    }
    // the iterate ends in either buffer: it becomes the solution buffer (both
    // hold zero ghosts and padding, every solve rewrites the interior)
    if (xa != xsol.p) std::swap(xsol.p, xalt.p);
    if (prof.enabled) {
        CATHODE_HIP(hipStreamSynchronize(s));
        for (size_t k = 0; k + 1 < prof.ev.size(); k += 2) {
            float ms = 0.f;
            CATHODE_HIP(hipEventElapsedTime(&ms, prof.ev[k], prof.ev[k + 1]));
            prof.total_ms += ms;
            prof.launches++;
        }
        for (auto& e : prof.ev) (void)hipEventDestroy(e);
        prof.ev.clear();
        // level-0 fused kernels (down and up alternate): per fine cell the down
        // kernel reads 5 coefficients + b and writes x and 1/4 coarse b (7.25
        // values), the up kernel also reads x and 1/4 coarse x (8.25 values)
        const double vb = mixed ? 4.0 : 8.0;
        // finest-level down kernel, per cell: aP aE aW aN aS and the activity
        // byte read, x and a quarter value of coarse b written; unfused, b
        // is read; fused, r and v (and the old p every other launch) are read
        // in FP64, the new p or s (FP64) and the level's b are written
        const double per_cell = fused ? 7.25 * vb + 1.0 + 8.0 * (2.5 + 1.0) : 7.25 * vb + 1.0;
        prof.bytes_per_launch = per_cell * (double)F.ni * (double)F.nj;
    }
    // a solve that ran no restart still leaves a whole hierarchy (a cut
    // solve leaves the chain to the next solve that needs it)
    if (cut < 0) flush_chain(s);
    // the cut the measured excess fractions allow for the next solve of
    // this kind (the restart reduced them, the x update copied them out)
    if (pulled && dpart_ok && !aborted) {
        if (cut_valid(0) && h_scal[SC_D0] >= CUT_DELTA0) cut_nx = 0;
        else if (setup_dcl > 0 && cut_valid(setup_dcl) && h_scal[SC_DC] >= CUT_DELTAC) cut_nx = setup_dcl;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->iterations = it;
        st->rel_residual = rel;
        st->converged = (rel <= tol) ? 1 : 0;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return it;
}

// ------------------------------------------------------------------ partial pre-reduction
// Fused kernels whose every workgroup needs a global sum (the BiCGStab
// scalars, the warm start's products, the setup's max|aP| and |b|^2) read
// the previous launch's per-workgroup partials.  That read is
// (workgroups) x (partials), so above presum_min() partials one workgroup sums
// each vector first, in a fixed order, and the consumers read one value.
constexpr int PRESUM_NT = 1024;
static std::atomic<int> g_presum_min{PRESUM_MIN_DEFAULT};
int presum_min() { return g_presum_min.load(std::memory_order_relaxed); }
void set_presum_min(int n) { g_presum_min.store(n < 0 ? PRESUM_MIN_DEFAULT : n, std::memory_order_relaxed); }

// dst[v] = sum (max where bit v of max_mask is set) of src[v * n + q], q < n
__global__ void __launch_bounds__(PRESUM_NT) k_presum(const double* __restrict__ src, int n, int nvec,
                                                      unsigned max_mask, double* __restrict__ dst)
{
    __shared__ double lds[PRESUM_NT / 64];
    for (int v = 0; v < nvec; v++) {
        const bool mx = (max_mask >> v) & 1u;
        const double* p = src + (size_t)v * n;
        double a = mx ? -INFINITY : 0.0;
        for (int q = threadIdx.x; q < n; q += PRESUM_NT) a = mx ? fmax(a, p[q]) : a + p[q];
        a = mx ? block_max<PRESUM_NT>(a, lds) : block_sum<PRESUM_NT>(a, lds);
        if (threadIdx.x == 0) dst[v] = a;
    }
}

// the (pointer, count) a consumer of nvec partial vectors of length n
// (vector v at src + v * n) reads: src and n, or the pre-reduced slots and 1
const double* MGSolver::presummed(const double* src, int n, int nvec, unsigned max_mask, int slot, int& count,
                                  hipStream_t s)
{
    count = n;
    if (n <= presum_min()) return src;
    presum_launch(src, n, nvec, max_mask, presum.p + slot, s);
    count = 1;
    return presum.p + slot;
}

void presum_launch(const double* src, int n, int nvec, unsigned max_mask, double* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_presum, dim3(1), dim3(PRESUM_NT), 0, s, src, n, nvec, max_mask, dst);
    CATHODE_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ reductions (common.hpp)
__global__ void k_reduce_partials(const double* part, int nblocks, int nvec, const int* is_max,
                                  double* out)
{
    __shared__ double lds[8];
    for (int v = 0; v < nvec; v++) {
        const bool mx = is_max && is_max[v];
        double s = mx ? -INFINITY : 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
            double p = part[(size_t)b * nvec + v];
            s = mx ? fmax(s, p) : s + p;
        }
        double t = mx ? block_max<256>(s, lds) : block_sum<256>(s, lds);
        if (threadIdx.x == 0) out[v] = t;
    }
}

__global__ void k_fetch_host(const double* __restrict__ d, size_t n, double* __restrict__ h)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) h[k] = d[k];
}

void fetch_host(double* h, const double* d, size_t n, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fetch_host, dim3(std::min(cdiv((long)n, 256), 16)), dim3(256), 0, s, d, n, h);
    CATHODE_LAUNCH_CHECK();
}

void reduce_partials(const double* part, int nblocks, int nvec, const int* d_is_max, double* out,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, part, nblocks, nvec, d_is_max, out);
    CATHODE_LAUNCH_CHECK();
}

}  // namespace cathode
