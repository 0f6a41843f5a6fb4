// vcycle.hpp -- fused V-cycle kernels (included by mg.hip).
//
// Fine levels run two kernels each:
//   k_down  a TI x TJ tile of cells plus a halo of 5: coefficients held in
//           registers, x in LDS, starting from x = 0; two red-black
//           Gauss-Seidel sweeps (4 half sweeps, red first), then the
//           residual of the tile cells, restricted (w-weighted sum of the
//           2x2 children) into the coarse right-hand side.  Each half sweep
//           invalidates one more ring of the halo, so after 4 sweeps and the
//           residual (5 rings) the tile holds exactly the values the global
//           red-black sweep produces.
//   k_up    the same tile with a halo of 4: x = x_down + x_coarse(parent)
//           (piecewise-constant prolongation), then two red-black sweeps.
//           In a V-cycle cut at this level (no coarser level visited) the
//           coarse array is null and x = x_down.
// Coarse levels (<= TAIL_MAX_CELLS cells each) run in one workgroup, all in
// LDS (k_tail), down to a single coarsest cell solved by one division.
#pragma once

#include <type_traits>

#include "vfuse.hpp"

namespace cathode {

// Tile shapes of the fused fine-level kernels: TI x TJ cells per workgroup of
// NT threads.  The multigrid picks one per level (autotuned on the first solve
// of a shape); index 0 is the default.
struct TileCfg {
    int ti, tj, nt;
};
constexpr int VT_NCFG = 10;
constexpr TileCfg VT_CFGS[VT_NCFG] = {{16, 32, 256}, {16, 32, 512}, {8, 32, 256},   {16, 64, 512},
                                      {32, 32, 512}, {8, 64, 512},  {8, 16, 256},   {16, 16, 256},
                                      {32, 64, 1024}, {64, 32, 1024}};  // the last two: less halo on large grids
constexpr int TAIL_MAX_CELLS = 512;     // levels with at most this many cells go to the tail
constexpr int TAIL_NT = 512;            // threads of the tail workgroup
constexpr int TAIL_MAX_LEVELS = 16;
constexpr int COARSEST_MAX = 1;         // complex path: coarsest level of one cell (one division)

// XCD-aware tile order.  Workgroups are dealt to the 8 XCDs round-robin
// (linear id b goes to XCD b % 8), each XCD with its own L2.  Mapping the
// tiles of XCD x to one contiguous run of the raster order keeps the halo
// rows neighbouring tiles share inside one L2.  Bijective for any count.
constexpr int N_XCD = 8;
__device__ inline void xcd_tile(int& ti, int& tj)
{
    const int n = (int)(gridDim.x * gridDim.y);
    const int b = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    const int q = n / N_XCD, r = n % N_XCD, x = b % N_XCD;
    const int t = x * q + min(x, r) + b / N_XCD;
    ti = t / (int)gridDim.x;
    tj = t - ti * (int)gridDim.x;
}

// b, x: the level's right-hand side and the down kernel's output (the
// pre-smoothed x).  xo: where the up kernel writes the corrected,
// post-smoothed x.  The up kernel reads x over its whole halo region, so it
// must never write x itself: a neighbouring tile that starts later would
// read post-smoothed halo values instead of the pre-smoothed ones the tile
// decomposition is exact for.
template <class T>
struct VLevel {
    int ni, nj, ld;
    const T *aP, *aE, *aW, *aN, *aS;
    const unsigned char* act;
    T *b, *x, *xo;
};

// ------------------------------------------------------------------ fine-level tiles
// TOP: the finest level (its own instantiation, so profiles separate it;
// only there are the restriction weights w applied).
template <class T, int TI, int TJ, int NT, int H, bool DOWN, bool TOP>
__global__ void __launch_bounds__(NT)
    k_vtile(VLevel<T> F, int fi, int fj, const double* __restrict__ w, int cni, int cnj, int cld,
            T* __restrict__ cbx, VFuse fu)
{
    constexpr bool REAL = std::is_floating_point<T>::value;
    constexpr int RI = TI + 2 * H, RJ = TJ + 2 * H, R = RI * RJ;
    constexpr int K = (R + NT - 1) / NT;
    __shared__ T xs[R];
    __shared__ T rs[DOWN ? TI * TJ : 1];
    const int tid = threadIdx.x;
    int bi, bj;  // this workgroup's tile
    xcd_tile(bi, bj);
    const int gi0 = bi * TI + 1 - H;  // global i of region row 0
    const int gj0 = bj * TJ + 1 - H;
    T cP[K], cE[K], cW[K], cN[K], cS[K], cB[K];
    unsigned flags[K];  // bit0 active, bit1 colour, bit2 updatable (not on the region rim), bit3 tile cell
    // fused BiCGStab inputs of the down prologue (FP64 r, v and the old p),
    // loaded before the scalars that combine them are known
    double fr[K], fv[K], fp[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int c = tid + k * NT;
        flags[k] = 0;
        cP[k] = T(1);
        cE[k] = cW[k] = cN[k] = cS[k] = cB[k] = T(0);
        if (c >= R) continue;
        const int ri = c / RJ, rj = c - ri * RJ;
        const int gi = gi0 + ri, gj = gj0 + rj;
        T xv = T(0);
        if (gi >= 1 && gi <= F.ni && gj >= 1 && gj <= F.nj) {
            const size_t g = (size_t)gi * F.ld + gj;
            if (F.act[g]) {
                flags[k] |= 1u;
                cP[k] = F.aP[g];
                cE[k] = F.aE[g];
                cW[k] = F.aW[g];
                cN[k] = F.aN[g];
                cS[k] = F.aS[g];
                if constexpr (REAL && DOWN && TOP) {
                    if (fu.pro) {
                        fr[k] = fu.r[g];
                        fv[k] = (fu.pro == 2 || !fu.first) ? fu.v[g] : 0.0;
                        fp[k] = (fu.pro == 1 && !fu.first) ? fu.pin[g] : 0.0;
                    } else {
                        cB[k] = F.b[g];
                    }
                } else {
                    cB[k] = F.b[g];
                }
                if (!DOWN) {
                    // no coarse level below (a hierarchy cut at this level): no correction
                    const int I = fi == 2 ? (gi + 1) >> 1 : gi, J = fj == 2 ? (gj + 1) >> 1 : gj;
                    xv = cbx ? F.x[g] + cbx[(size_t)I * cld + J] : F.x[g];
                }
            }
        }
        flags[k] |= ((gi + gj) & 1) ? 2u : 0u;
        if (ri >= 1 && ri <= RI - 2 && rj >= 1 && rj <= RJ - 2) flags[k] |= 4u;
        if (ri >= H && ri < H + TI && rj >= H && rj < H + TJ && gi <= F.ni && gj <= F.nj) flags[k] |= 8u;
        xs[c] = xv;
    }
    // fused BiCGStab scalars of the down prologue: q = r - al v (pro 2) or
    // r + beta (p_old - om v) (pro 1); f scales the FP32 input
    if constexpr (REAL && DOWN && TOP) {
        if (fu.pro) {
            double f_al = 0.0, f_beta = 0.0, f_om = 0.0, f_in = 1.0;
            const bool lead = tid == 0 && blockIdx.x == 0 && blockIdx.y == 0;
            if (fu.pred) {
                __shared__ double pr[NT / 64];
                double a = 0.0;
                for (int k = tid; k < fu.npred; k += NT) a += fu.pred[k];
                a = block_sum_bcast<NT>(a, pr);
                if (fu.pro == 2) {
                    const double al = fu.sc[fu.s_rho1] / a;
                    f_al = (a != 0.0 && isfinite(al)) ? al : 0.0;  // breakdown: no step
                    if (lead) {
                        fu.sc[fu.s_r0v] = a;
                        fu.sc[fu.s_al] = f_al;
                    }
                } else {
                    if (!fu.first) f_beta = bicg_beta(a, fu.sc[fu.s_rho], fu.sc[fu.s_al], fu.sc[fu.s_om]);
                    if (lead) fu.sc[fu.s_rho1] = a;
                }
            } else if (fu.pro == 2) {
                f_al = fu.sc[fu.s_al];
            } else if (!fu.first) {
                f_beta = bicg_beta(fu.sc[fu.s_rho1], fu.sc[fu.s_rho], fu.sc[fu.s_al], fu.sc[fu.s_om]);
            }
            f_om = fu.sc[fu.s_om];
            f_in = fu.mixed ? fu.sc[fu.s_s32] * fu.sc[fu.s_scale] : 1.0;
            // q = r - al v (pro 2), r (first), r + beta (p_old - om v) (pro 1)
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (!(flags[k] & 1u)) continue;
                double q;
                if (fu.pro == 2) {
                    q = fr[k] - f_al * fv[k];
                } else if (fu.first) {
                    q = fr[k];
                } else {
                    q = fr[k] + f_beta * (fp[k] - f_om * fv[k]);
                }
                cB[k] = (T)(q * f_in);
                if (flags[k] & 8u) {  // the old p is read from pin, so no tile reads what another writes
                    const int c = tid + k * NT, ri = c / RJ, rj = c - ri * RJ;
                    const size_t g = (size_t)(gi0 + ri) * F.ld + (gj0 + rj);
                    fu.ps[g] = q;
                    F.b[g] = cB[k];
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const unsigned col = (unsigned)(s & 1) << 1;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int c = tid + k * NT;
            if ((flags[k] & 7u) == (5u | col)) {
                T v = cB[k] + cE[k] * xs[c + RJ] + cW[k] * xs[c - RJ] + cN[k] * xs[c + 1] + cS[k] * xs[c - 1];
                xs[c] = v / cP[k];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int c = tid + k * NT;
        if (!(flags[k] & 8u)) continue;
        const int ri = c / RJ, rj = c - ri * RJ;
        const int gi = gi0 + ri, gj = gj0 + rj;
        (DOWN ? F.x : F.xo)[(size_t)gi * F.ld + gj] = xs[c];
        if (DOWN) {
            T r = T(0);
            if (flags[k] & 1u) {
                r = cB[k] - cP[k] * xs[c] + cE[k] * xs[c + RJ] + cW[k] * xs[c - RJ] + cN[k] * xs[c + 1] +
                    cS[k] * xs[c - 1];
                if (TOP && w) r = r * (T)w[gj];
            }
            rs[(ri - H) * TJ + (rj - H)] = r;
        }
    }
    if (DOWN) {
        __syncthreads();
        // coarse cells of this tile
        const int CI = TI / fi, CJ = TJ / fj;
        const int I0 = bi * CI, J0 = bj * CJ;
        for (int q = tid; q < CI * CJ; q += NT) {
            const int il = q / CJ, jl = q - il * CJ;
            const int I = I0 + il + 1, J = J0 + jl + 1;
            if (I > cni || J > cnj) continue;
            T sum = T(0);
            for (int a = 0; a < fi; a++)
                for (int b = 0; b < fj; b++) {
                    const int ti = il * fi + a, tj = jl * fj + b;
                    const int gi = bi * TI + ti + 1, gj = bj * TJ + tj + 1;
                    if (gi <= F.ni && gj <= F.nj) sum += rs[ti * TJ + tj];
                }
            cbx[(size_t)I * cld + J] = sum;
        }
    }
    if constexpr (REAL && !DOWN && TOP) {
        if (fu.epi) {
            // y = V-cycle output in FP64, out = A y (FP64 finest operator),
            // partial dot products of this tile
            __shared__ double red[NT / 64 > 8 ? NT / 64 : 8];
            const double oscale = fu.mixed ? 1.0 / fu.sc[fu.s_scale] : 1.0;
            double d1 = 0.0, d2 = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int c = tid + k * NT;
                if (!(flags[k] & 8u)) continue;
                const int ri = c / RJ, rj = c - ri * RJ;
                const size_t g = (size_t)(gi0 + ri) * F.ld + (gj0 + rj);
                double yk = 0.0, o = 0.0;
                if (flags[k] & 1u) {
                    yk = (double)xs[c] * oscale;
                    o = fu.aP[g] * yk - fu.aE[g] * ((double)xs[c + RJ] * oscale) -
                        fu.aW[g] * ((double)xs[c - RJ] * oscale) - fu.aN[g] * ((double)xs[c + 1] * oscale) -
                        fu.aS[g] * ((double)xs[c - 1] * oscale);
                    d1 = d1 + fu.dotvec[g] * o;
                    d2 = d2 + o * o;
                }
                fu.y[g] = yk;
                fu.out[g] = o;
            }
            const double t1 = block_sum<NT>(d1, red);
            const double t2 = fu.want_tt ? block_sum<NT>(d2, red) : 0.0;
            if (tid == 0) {
                const int ntiles = (int)(gridDim.x * gridDim.y);
                const int tile = bi * (int)gridDim.x + bj;
                fu.part[tile] = t1;
                if (fu.want_tt) fu.part[ntiles + tile] = t2;
            }
        }
    }
}

// Launches the down (DOWN) or up kernel of one fine level with tile config cfg.
// With ev0/ev1 set the launch goes through hipExtLaunchKernelGGL, whose
// events carry the dispatch's own start / end timestamps (kernel time only,
// as a profiler reports it).
// The finest level's up kernel uses a halo of 5 (the fused epilogue needs
// the ring next to the tile); the other up kernels 4, the down kernels 5.
template <class T, bool DOWN, int C>
inline void launch_vtile_cfg(bool top, hipStream_t s, const VLevel<T>& V, int fi, int fj, const double* w, int cni,
                             int cnj, int cld, T* cbx, const VFuse& fu = VFuse(), hipEvent_t ev0 = nullptr,
                             hipEvent_t ev1 = nullptr)
{
    constexpr TileCfg G = VT_CFGS[C];
    constexpr int HN = DOWN ? 5 : 4;  // not the finest level
    const dim3 grid(cdiv(V.nj, G.tj), cdiv(V.ni, G.ti));
    if (ev0) {
        if (top)
            hipExtLaunchKernelGGL((k_vtile<T, G.ti, G.tj, G.nt, 5, DOWN, true>), grid, dim3(G.nt), 0, s, ev0, ev1, 0,
                                  V, fi, fj, w, cni, cnj, cld, cbx, fu);
        else
            hipExtLaunchKernelGGL((k_vtile<T, G.ti, G.tj, G.nt, HN, DOWN, false>), grid, dim3(G.nt), 0, s, ev0, ev1,
                                  0, V, fi, fj, w, cni, cnj, cld, cbx, fu);
    } else if (top) {
        hipLaunchKernelGGL((k_vtile<T, G.ti, G.tj, G.nt, 5, DOWN, true>), grid, dim3(G.nt), 0, s, V, fi, fj, w, cni,
                           cnj, cld, cbx, fu);
    } else {
        hipLaunchKernelGGL((k_vtile<T, G.ti, G.tj, G.nt, HN, DOWN, false>), grid, dim3(G.nt), 0, s, V, fi, fj, w, cni,
                           cnj, cld, cbx, fu);
    }
}

template <class T, bool DOWN>
inline void launch_vtile(int cfg, bool top, hipStream_t s, const VLevel<T>& V, int fi, int fj, const double* w,
                         int cni, int cnj, int cld, T* cbx, const VFuse& fu = VFuse(), hipEvent_t ev0 = nullptr,
                         hipEvent_t ev1 = nullptr)
{
    switch (cfg) {
    case 1: launch_vtile_cfg<T, DOWN, 1>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    case 2: launch_vtile_cfg<T, DOWN, 2>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    case 3: launch_vtile_cfg<T, DOWN, 3>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    case 4: launch_vtile_cfg<T, DOWN, 4>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    case 5: launch_vtile_cfg<T, DOWN, 5>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    case 6: launch_vtile_cfg<T, DOWN, 6>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    case 7: launch_vtile_cfg<T, DOWN, 7>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    case 8: launch_vtile_cfg<T, DOWN, 8>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    case 9: launch_vtile_cfg<T, DOWN, 9>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    default: launch_vtile_cfg<T, DOWN, 0>(top, s, V, fi, fj, w, cni, cnj, cld, cbx, fu, ev0, ev1); break;
    }
}

// tiles of the finest level with configuration cfg (partials of the fused epilogue)
inline int vtile_count(int cfg, int ni, int nj)
{
    const TileCfg G = VT_CFGS[cfg < 0 || cfg >= VT_NCFG ? 0 : cfg];
    return cdiv(nj, G.tj) * cdiv(ni, G.ti);
}

// ------------------------------------------------------------------ two fine levels per launch
// A fine level F that coarsens 2x2 into the next level C runs, with C, in one
// down and one up launch instead of two each.  A workgroup owns a TI x TJ
// tile of C and the 2TI x 2TJ tile of F it aggregates.  Every value it
// writes is the one the single-level kernels write (same operations in the
// same order), so the results are bitwise those of the single-level pair.
//
// Down (halo 5 per level): the C region is the C tile plus 5 rings; its b is
// the restriction of F's residuals on the children of the C region, which
// F's 2 sweeps leave exact on F's region shrunk by 5 rings, so F's region is
// those children plus 5 rings.  Writes F's x and C's b and x on the owned
// tiles and the restriction of C's residual into the level below C (cb).
//
// Up (halo 4 per level): F's region is its tile plus 4 rings; the C cells
// that are parents of it (C tile plus 2 rings) must hold their
// post-smoothed x, so C's region is those plus 4 rings.  C starts from its
// down x plus the correction from the level below (cx3); F from its down x
// plus C's post-smoothed x.  Writes F's xo on the owned tile (C's
// post-smoothed x has no reader outside the launch).

// one level's region in LDS: coefficients of the NT-strided cells in registers
template <class T, int RI, int RJ, int NT>
struct PairRegion {
    static constexpr int R = RI * RJ, K = (R + NT - 1) / NT;
    T cP[K], cE[K], cW[K], cN[K], cS[K], cB[K];
    unsigned flags[K];  // bit0 active, bit1 colour, bit2 updatable (not on the region rim)
};

// the coefficients of region cells (gi0, gj0: level index of region row /
// column 0); cells outside the level are inactive
template <class T, int RI, int RJ, int NT>
__device__ inline void pair_load(PairRegion<T, RI, RJ, NT>& G, const VLevel<T>& L, int gi0, int gj0)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < G.K; k++) {
        const int c = tid + k * NT;
        G.flags[k] = 0;
        G.cP[k] = T(1);
        G.cE[k] = G.cW[k] = G.cN[k] = G.cS[k] = G.cB[k] = T(0);
        if (c >= G.R) continue;
        const int ri = c / RJ, rj = c - ri * RJ;
        const int gi = gi0 + ri, gj = gj0 + rj;
        if (gi >= 1 && gi <= L.ni && gj >= 1 && gj <= L.nj) {
            const size_t g = (size_t)gi * L.ld + gj;
            if (L.act[g]) {
                G.flags[k] |= 1u;
                G.cP[k] = L.aP[g];
                G.cE[k] = L.aE[g];
                G.cW[k] = L.aW[g];
                G.cN[k] = L.aN[g];
                G.cS[k] = L.aS[g];
            }
        }
        G.flags[k] |= ((gi + gj) & 1) ? 2u : 0u;
        if (ri >= 1 && ri <= RI - 2 && rj >= 1 && rj <= RJ - 2) G.flags[k] |= 4u;
    }
}

// 4 half sweeps (red first) of the region in xs; ends with a barrier
template <class T, int RI, int RJ, int NT>
__device__ inline void pair_sweeps4(const PairRegion<T, RI, RJ, NT>& G, T* xs)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const unsigned col = (unsigned)(s & 1) << 1;
#pragma unroll
        for (int k = 0; k < G.K; k++) {
            const int c = tid + k * NT;
            if ((G.flags[k] & 7u) == (5u | col)) {
                T v = G.cB[k] + G.cE[k] * xs[c + RJ] + G.cW[k] * xs[c - RJ] + G.cN[k] * xs[c + 1] +
                      G.cS[k] * xs[c - 1];
                xs[c] = v / G.cP[k];
            }
        }
        __syncthreads();
    }
}

template <class T>
__device__ inline T pair_residual(T b, T P, T E, T W, T N, T S, const T* xs, int c, int RJ)
{
    return b - P * xs[c] + E * xs[c + RJ] + W * xs[c - RJ] + N * xs[c + 1] + S * xs[c - 1];
}

template <class T, int TI, int TJ, int NT>
__global__ void __launch_bounds__(NT)
    k_vpair_down(VLevel<T> F, VLevel<T> C, int fi3, int fj3, int c3ni, int c3nj, int c3ld, T* __restrict__ cb3)
{
    constexpr int H = 5;
    constexpr int CI = TI + 2 * H, CJ = TJ + 2 * H;      // C region
    constexpr int QI = 2 * CI, QJ = 2 * CJ;              // F cells restricted into it
    constexpr int FI = QI + 2 * H, FJ = QJ + 2 * H;      // F region
    __shared__ T xf[FI * FJ];
    __shared__ T rq[QI * QJ];
    __shared__ T xc[CI * CJ];
    __shared__ T rc[TI * TJ];
    const int tid = threadIdx.x;
    int bi, bj;  // this workgroup's C tile
    xcd_tile(bi, bj);
    const int ci0 = bi * TI + 1 - H, cj0 = bj * TJ + 1 - H;  // C index of C region row / column 0
    const int fi0 = 2 * ci0 - 1 - H, fj0 = 2 * cj0 - 1 - H;  // F index of F region row / column 0
    PairRegion<T, FI, FJ, NT> Gf;
    PairRegion<T, CI, CJ, NT> Gc;
    pair_load(Gf, F, fi0, fj0);
    pair_load(Gc, C, ci0, cj0);
#pragma unroll
    for (int k = 0; k < Gf.K; k++) {
        const int c = tid + k * NT;
        if (c >= Gf.R) continue;
        if (Gf.flags[k] & 1u) {
            const int ri = c / FJ, rj = c - ri * FJ;
            Gf.cB[k] = F.b[(size_t)(fi0 + ri) * F.ld + (fj0 + rj)];
        }
        xf[c] = T(0);
    }
    for (int c = tid; c < CI * CJ; c += NT) xc[c] = T(0);
    __syncthreads();
    pair_sweeps4(Gf, xf);
    // F's x on its tile, F's residuals on the children of the C region
#pragma unroll
    for (int k = 0; k < Gf.K; k++) {
        const int c = tid + k * NT;
        if (c >= Gf.R) continue;
        const int ri = c / FJ, rj = c - ri * FJ;
        const int gi = fi0 + ri, gj = fj0 + rj;
        if (ri >= 3 * H && ri < 3 * H + 2 * TI && rj >= 3 * H && rj < 3 * H + 2 * TJ && gi <= F.ni && gj <= F.nj)
            F.x[(size_t)gi * F.ld + gj] = xf[c];
        if (ri >= H && ri < H + QI && rj >= H && rj < H + QJ) {
            T r = T(0);
            if (Gf.flags[k] & 1u) r = pair_residual(Gf.cB[k], Gf.cP[k], Gf.cE[k], Gf.cW[k], Gf.cN[k], Gf.cS[k], xf, c, FJ);
            rq[(ri - H) * QJ + (rj - H)] = r;
        }
    }
    __syncthreads();
    // C's b: the sum of its children's residuals (children inside F)
#pragma unroll
    for (int k = 0; k < Gc.K; k++) {
        const int c = tid + k * NT;
        if (c >= Gc.R) continue;
        const int ri = c / CJ, rj = c - ri * CJ;
        const int gi = ci0 + ri, gj = cj0 + rj;
        if (gi < 1 || gi > C.ni || gj < 1 || gj > C.nj) continue;
        T sum = T(0);
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                const int fi = 2 * gi - 1 + a, fj = 2 * gj - 1 + b;
                if (fi <= F.ni && fj <= F.nj) sum += rq[(2 * ri + a) * QJ + (2 * rj + b)];
            }
        if (Gc.flags[k] & 1u) Gc.cB[k] = sum;
        if (ri >= H && ri < H + TI && rj >= H && rj < H + TJ) C.b[(size_t)gi * C.ld + gj] = sum;
    }
    pair_sweeps4(Gc, xc);
#pragma unroll
    for (int k = 0; k < Gc.K; k++) {
        const int c = tid + k * NT;
        if (c >= Gc.R) continue;
        const int ri = c / CJ, rj = c - ri * CJ;
        const int gi = ci0 + ri, gj = cj0 + rj;
        if (!(ri >= H && ri < H + TI && rj >= H && rj < H + TJ && gi <= C.ni && gj <= C.nj)) continue;
        C.x[(size_t)gi * C.ld + gj] = xc[c];
        T r = T(0);
        if (Gc.flags[k] & 1u) r = pair_residual(Gc.cB[k], Gc.cP[k], Gc.cE[k], Gc.cW[k], Gc.cN[k], Gc.cS[k], xc, c, CJ);
        rc[(ri - H) * TJ + (rj - H)] = r;
    }
    __syncthreads();
    // the level below C: coarse cells of this tile
    const int KI = TI / fi3, KJ = TJ / fj3;
    const int I0 = bi * KI, J0 = bj * KJ;
    for (int q = tid; q < KI * KJ; q += NT) {
        const int il = q / KJ, jl = q - il * KJ;
        const int I = I0 + il + 1, J = J0 + jl + 1;
        if (I > c3ni || J > c3nj) continue;
        T sum = T(0);
        for (int a = 0; a < fi3; a++)
            for (int b = 0; b < fj3; b++) {
                const int ti = il * fi3 + a, tj = jl * fj3 + b;
                const int gi = bi * TI + ti + 1, gj = bj * TJ + tj + 1;
                if (gi <= C.ni && gj <= C.nj) sum += rc[ti * TJ + tj];
            }
        cb3[(size_t)I * c3ld + J] = sum;
    }
}

template <class T, int TI, int TJ, int NT>
__global__ void __launch_bounds__(NT)
    k_vpair_up(VLevel<T> F, VLevel<T> C, int fi3, int fj3, int c3ld, const T* __restrict__ cx3)
{
    constexpr int H = 4;
    constexpr int FI = 2 * TI + 2 * H, FJ = 2 * TJ + 2 * H;  // F region
    constexpr int CI = TI + 2 + 2 * H + 2, CJ = TJ + 2 + 2 * H + 2;  // C region: parents of F's region + 4 rings
    __shared__ T xf[FI * FJ];
    __shared__ T xc[CI * CJ];
    const int tid = threadIdx.x;
    int bi, bj;
    xcd_tile(bi, bj);
    const int fi0 = 2 * bi * TI + 1 - H, fj0 = 2 * bj * TJ + 1 - H;  // F index of F region row / column 0
    const int ci0 = bi * TI + 1 - 2 - H, cj0 = bj * TJ + 1 - 2 - H;  // C index of C region row / column 0
    PairRegion<T, FI, FJ, NT> Gf;
    PairRegion<T, CI, CJ, NT> Gc;
    pair_load(Gf, F, fi0, fj0);
    pair_load(Gc, C, ci0, cj0);
    T xdf[Gf.K];  // F's down x, corrected once C is post-smoothed
#pragma unroll
    for (int k = 0; k < Gf.K; k++) {
        xdf[k] = T(0);
        const int c = tid + k * NT;
        if (c >= Gf.R || !(Gf.flags[k] & 1u)) continue;
        const int ri = c / FJ, rj = c - ri * FJ;
        const size_t g = (size_t)(fi0 + ri) * F.ld + (fj0 + rj);
        Gf.cB[k] = F.b[g];
        xdf[k] = F.x[g];
    }
#pragma unroll
    for (int k = 0; k < Gc.K; k++) {
        const int c = tid + k * NT;
        if (c >= Gc.R) continue;
        T xv = T(0);
        if (Gc.flags[k] & 1u) {
            const int ri = c / CJ, rj = c - ri * CJ;
            const int gi = ci0 + ri, gj = cj0 + rj;
            const size_t g = (size_t)gi * C.ld + gj;
            Gc.cB[k] = C.b[g];
            const int I = fi3 == 2 ? (gi + 1) >> 1 : gi, J = fj3 == 2 ? (gj + 1) >> 1 : gj;
            xv = cx3 ? C.x[g] + cx3[(size_t)I * c3ld + J] : C.x[g];  // cx3 null: C is the last level
        }
        xc[c] = xv;
    }
    __syncthreads();
    pair_sweeps4(Gc, xc);
    // F: down x plus the parent's post-smoothed x
#pragma unroll
    for (int k = 0; k < Gf.K; k++) {
        const int c = tid + k * NT;
        if (c >= Gf.R) continue;
        T xv = T(0);
        if (Gf.flags[k] & 1u) {
            const int ri = c / FJ, rj = c - ri * FJ;
            const int gi = fi0 + ri, gj = fj0 + rj;
            const int pi = ((gi + 1) >> 1) - ci0, pj = ((gj + 1) >> 1) - cj0;
            xv = xdf[k] + xc[pi * CJ + pj];
        }
        xf[c] = xv;
    }
    __syncthreads();
    pair_sweeps4(Gf, xf);
#pragma unroll
    for (int k = 0; k < Gf.K; k++) {
        const int c = tid + k * NT;
        if (c >= Gf.R) continue;
        const int ri = c / FJ, rj = c - ri * FJ;
        const int gi = fi0 + ri, gj = fj0 + rj;
        if (ri >= H && ri < H + 2 * TI && rj >= H && rj < H + 2 * TJ && gi <= F.ni && gj <= F.nj)
            F.xo[(size_t)gi * F.ld + gj] = xf[c];
    }
}

// the pair kernels' tiles of the second level and workgroup sizes
constexpr int VP_NCFG = 6;
constexpr TileCfg VP_CFGS[VP_NCFG] = {{16, 16, 1024}, {8, 16, 512}, {8, 16, 1024}, {8, 8, 1024}, {8, 8, 512},
                                      {4, 8, 1024}};

template <class T, int CFG>
inline void launch_vpair_cfg(bool down, hipStream_t s, const VLevel<T>& F, const VLevel<T>& C, int fi3, int fj3,
                             int c3ni, int c3nj, int c3ld, T* c3)
{
    constexpr TileCfg G = VP_CFGS[CFG];
    const dim3 grid(cdiv(C.nj, G.tj), cdiv(C.ni, G.ti));
    if (down)
        hipLaunchKernelGGL((k_vpair_down<T, G.ti, G.tj, G.nt>), grid, dim3(G.nt), 0, s, F, C, fi3, fj3, c3ni, c3nj,
                           c3ld, c3);
    else
        hipLaunchKernelGGL((k_vpair_up<T, G.ti, G.tj, G.nt>), grid, dim3(G.nt), 0, s, F, C, fi3, fj3, c3ld,
                           (const T*)c3);
}

template <class T>
inline void launch_vpair(int cfg, bool down, hipStream_t s, const VLevel<T>& F, const VLevel<T>& C, int fi3, int fj3,
                         int c3ni, int c3nj, int c3ld, T* c3)
{
    switch (cfg) {
    case 1: launch_vpair_cfg<T, 1>(down, s, F, C, fi3, fj3, c3ni, c3nj, c3ld, c3); break;
    case 2: launch_vpair_cfg<T, 2>(down, s, F, C, fi3, fj3, c3ni, c3nj, c3ld, c3); break;
    case 3: launch_vpair_cfg<T, 3>(down, s, F, C, fi3, fj3, c3ni, c3nj, c3ld, c3); break;
    case 4: launch_vpair_cfg<T, 4>(down, s, F, C, fi3, fj3, c3ni, c3nj, c3ld, c3); break;
    case 5: launch_vpair_cfg<T, 5>(down, s, F, C, fi3, fj3, c3ni, c3nj, c3ld, c3); break;
    default: launch_vpair_cfg<T, 0>(down, s, F, C, fi3, fj3, c3ni, c3nj, c3ld, c3); break;
    }
}

// ------------------------------------------------------------------ coarse levels in one workgroup
template <class T>
struct TailArgs {
    int nt;                          // number of tail levels (the last one is the coarsest)
    int ni[TAIL_MAX_LEVELS], nj[TAIL_MAX_LEVELS], off[TAIL_MAX_LEVELS];  // LDS cell offset per level
    int cx[TAIL_MAX_LEVELS], cy[TAIL_MAX_LEVELS];
    int gld[TAIL_MAX_LEVELS];        // global row pitch per level
    const T* aP[TAIL_MAX_LEVELS];
    const T* aE[TAIL_MAX_LEVELS];
    const T* aW[TAIL_MAX_LEVELS];
    const T* aN[TAIL_MAX_LEVELS];
    const T* aS[TAIL_MAX_LEVELS];
    const unsigned char* act[TAIL_MAX_LEVELS];
    int ncell;                       // total LDS cells
    const T* b_in;                   // right-hand side of the first tail level (global)
    T* x_out;                        // its solution (global)
    const double* w;                 // restriction weights when the tail starts at level 0
    const double* cinv;              // k_tail_reg: dense FP64 inverse of the coarsest level (cn x cn, row-major)
    const double* s32;               // FP32 levels: their scale (they hold s32 A); nullptr for FP64 levels
    int cn;                          // its cells
};

// two red-black sweeps (red first) of one tail level; c = this thread's cell
// (-1: none), colour its parity
template <class T>
__device__ inline void tail_sweeps(const T* P, const T* E, const T* W, const T* N, const T* S, const T* B, T* X,
                                   int ld, int c, int colour)
{
    for (int sw = 0; sw < 4; sw++) {
        if (c >= 0 && colour == (sw & 1))
            X[c] = (B[c] + E[c] * X[c + ld] + W[c] * X[c - ld] + N[c] * X[c + 1] + S[c] * X[c - 1]) / P[c];
        __syncthreads();
    }
}

template <class T>
__global__ void __launch_bounds__(TAIL_NT) k_tail(TailArgs<T> A)
{
    extern __shared__ unsigned char tail_lds[];
    T* base = reinterpret_cast<T*>(tail_lds);
    const int nc = A.ncell;
    T* P = base;
    T* E = base + nc;
    T* W = base + 2 * nc;
    T* N = base + 3 * nc;
    T* S = base + 4 * nc;
    T* B = base + 5 * nc;
    T* X = base + 6 * nc;
    unsigned char* ACT = reinterpret_cast<unsigned char*>(base + 7 * nc);
    // load every tail level (ghost ring: identity rows, x = b = 0)
    for (int l = 0; l < A.nt; l++) {
        const int ni = A.ni[l], nj = A.nj[l], ld = nj + 2, o = A.off[l];
        for (int q = threadIdx.x; q < (ni + 2) * ld; q += TAIL_NT) {
            const int i = q / ld, j = q % ld;
            const int c = o + q;
            bool a = false;
            if (i >= 1 && i <= ni && j >= 1 && j <= nj) {
                const size_t g = (size_t)i * A.gld[l] + j;
                a = A.act[l][g] != 0;
                if (a) {
                    P[c] = A.aP[l][g];
                    E[c] = A.aE[l][g];
                    W[c] = A.aW[l][g];
                    N[c] = A.aN[l][g];
                    S[c] = A.aS[l][g];
                }
                B[c] = (l == 0 && a) ? A.b_in[g] : T(0);
            } else {
                B[c] = T(0);
            }
            if (!a) {
                P[c] = T(1);
                E[c] = W[c] = N[c] = S[c] = T(0);
            }
            X[c] = T(0);
            ACT[c] = a ? 1 : 0;
        }
    }
    __syncthreads();
    // each thread owns cell threadIdx.x of every tail level (<= TAIL_NT cells)
    // down
    for (int l = 0; l + 1 < A.nt; l++) {
        const int ni = A.ni[l], nj = A.nj[l], ld = nj + 2, o = A.off[l];
        const int q = threadIdx.x;
        const int i = 1 + q / nj, j = 1 + q % nj;
        const int c = q < ni * nj ? o + i * ld + j : -1;
        tail_sweeps(P, E, W, N, S, B, X, ld, c, (i + j) & 1);
        // restriction: coarse cell q of level l+1
        const int fi = A.cx[l] ? 2 : 1, fj = A.cy[l] ? 2 : 1;
        const int cni = A.ni[l + 1], cnj = A.nj[l + 1], cld = cnj + 2, co = A.off[l + 1];
        if (q < cni * cnj) {
            const int I = 1 + q / cnj, J = 1 + q % cnj;
            T sum = T(0);
            for (int a = 0; a < fi; a++)
                for (int bb = 0; bb < fj; bb++) {
                    const int fi_ = fi * (I - 1) + 1 + a, fj_ = fj * (J - 1) + 1 + bb;
                    if (fi_ > ni || fj_ > nj) continue;
                    const int cc = o + fi_ * ld + fj_;
                    T r = B[cc] - P[cc] * X[cc] + E[cc] * X[cc + ld] + W[cc] * X[cc - ld] + N[cc] * X[cc + 1] +
                          S[cc] * X[cc - 1];
                    if (l == 0 && A.w) r = r * (T)A.w[fj_];
                    sum += r;
                }
            B[co + I * cld + J] = sum;
        }
        __syncthreads();
    }
    // coarsest level: a single cell, x = b / aP (inactive or zero diagonal: 0)
    {
        const int l = A.nt - 1;
        const int ni = A.ni[l], nj = A.nj[l], ld = nj + 2, o = A.off[l];
        const int q = threadIdx.x;
        if (q < ni * nj) {
            const int c = o + (1 + q / nj) * ld + 1 + q % nj;
            X[c] = (ACT[c] && !(P[c] == T(0))) ? B[c] / P[c] : T(0);
        }
        __syncthreads();
    }
    // up
    for (int l = A.nt - 2; l >= 0; l--) {
        const int ni = A.ni[l], nj = A.nj[l], ld = nj + 2, o = A.off[l];
        const int cld = A.nj[l + 1] + 2, co = A.off[l + 1];
        const int q = threadIdx.x;
        const int i = 1 + q / nj, j = 1 + q % nj;
        const int c = q < ni * nj ? o + i * ld + j : -1;
        if (c >= 0 && ACT[c]) {
            const int I = A.cx[l] ? (i + 1) >> 1 : i, J = A.cy[l] ? (j + 1) >> 1 : j;
            X[c] = X[c] + X[co + I * cld + J];
        }
        __syncthreads();
        tail_sweeps(P, E, W, N, S, B, X, ld, c, (i + j) & 1);
    }
    // write the first tail level
    {
        const int ni = A.ni[0], nj = A.nj[0], ld = nj + 2;
        for (int q = threadIdx.x; q < ni * nj; q += TAIL_NT) {
            const int i = 1 + q / nj, j = 1 + q % nj;
            A.x_out[(size_t)i * A.gld[0] + j] = X[i * ld + j];
        }
    }
}

// ------------------------------------------------------------------ register-resident tail (real types)
// Same levels, same sweeps and transfers as k_tail, organised for latency.
// The workgroup has TAILR_NT threads.  The first tail level may hold up to
// TAIL_TOP_MAX cells: thread q owns its C0 consecutive cells q*C0 ..
// q*C0+C0-1 (consecutive cells alternate in colour, so every half sweep
// keeps every thread busy); every further level has at most TAILR_NT cells,
// one per thread.  Each owned cell's coefficients and 1/aP sit in registers,
// LDS holds only x, b and the residuals, and phases whose cells all lie in
// the first wavefront (<= 64 cells) synchronise that wavefront alone instead
// of the whole workgroup.
constexpr int TAILR_NT = 1024;
constexpr int TAIL_REG_LEVELS = 5;    // the first tail level and at most 4 below it (register budget of the FP64 tail)
constexpr int TAIL_TOP_MAX = 2 * TAILR_NT;

__device__ inline void tail_sync(int cells)
{
    if (cells > 64) {
        __syncthreads();
    } else {
        // one wavefront: its LDS operations complete in issue order; keep
        // the compiler from moving them across the phase boundary
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// the cells of one tail level owned by one thread
template <class T, int C>
struct TailCells {
    T P[C], iP[C], E[C], W[C], N[C], S[C], wq[C];
    int c[C];     // LDS index, -1 = none
    int idx[C];   // cell number in the level (row-major over the interior)
    int par[C];   // LDS index of the parent in the next level
    unsigned fl[C];  // bit0 active, bit1 colour
};

// Two phases, so that every global load of the launch is in flight at once
// (a use of any loaded value waits for all loads issued before it):
// tail_fetch issues one level's loads -- coefficients, activity byte and, on
// the first level, b -- without using them (a thread without a cell reads
// the level's first cell and discards it); tail_finish applies the activity
// mask, forms 1/aP and puts the first level's b into LDS.
template <class T, int C>
__device__ inline void tail_fetch(TailCells<T, C>& L, const TailArgs<T>& A, int l, int q, T (&b)[C],
                                  unsigned (&act)[C])
{
    const int ni = A.ni[l], nj = A.nj[l], ld = nj + 2;
#pragma unroll
    for (int k = 0; k < C; k++) {
        const int ci = q * C + k;
        const bool here = ci < ni * nj;
        const int cc = here ? ci : 0;
        const int i = 1 + cc / nj, j = 1 + cc % nj;
        const size_t g = (size_t)i * A.gld[l] + j;
        act[k] = A.act[l][g];
        L.P[k] = A.aP[l][g];
        L.E[k] = A.aE[l][g];
        L.W[k] = A.aW[l][g];
        L.N[k] = A.aN[l][g];
        L.S[k] = A.aS[l][g];
        b[k] = l == 0 ? A.b_in[g] : T(0);
        L.wq[k] = (l == 0 && A.w) ? (T)A.w[j] : T(1);
        L.c[k] = here ? A.off[l] + i * ld + j : -1;
        L.idx[k] = here ? ci : 0;
        L.fl[k] = (here && ((i + j) & 1)) ? 2u : 0u;
        L.par[k] = 0;
        if (here && l + 1 < A.nt) {
            const int I = A.cx[l] ? (i + 1) >> 1 : i, J = A.cy[l] ? (j + 1) >> 1 : j;
            L.par[k] = A.off[l + 1] + I * (A.nj[l + 1] + 2) + J;
        }
    }
}

template <class T, int C>
__device__ inline void tail_finish(TailCells<T, C>& L, int l, const T (&b)[C], const unsigned (&act)[C], T* B)
{
#pragma unroll
    for (int k = 0; k < C; k++) {
        const bool on = L.c[k] >= 0 && act[k] != 0;
        L.fl[k] |= on ? 1u : 0u;
        L.iP[k] = on ? T(1) / L.P[k] : T(1);
        L.P[k] = on ? L.P[k] : T(1);
        L.E[k] = on ? L.E[k] : T(0);
        L.W[k] = on ? L.W[k] : T(0);
        L.N[k] = on ? L.N[k] : T(0);
        L.S[k] = on ? L.S[k] : T(0);
        if (l == 0 && on) B[L.c[k]] = b[k];
    }
}

// 4 half sweeps (red first); the last one is not followed by a barrier
template <class T, int C>
__device__ inline void tail_sweeps4(const TailCells<T, C>& L, T* X, const T* B, int ld, int cells)
{
    for (int s = 0; s < 4; s++) {
#pragma unroll
        for (int k = 0; k < C; k++) {
            const int c = L.c[k];
            if (c >= 0 && (L.fl[k] & 3u) == (1u | ((unsigned)(s & 1) << 1)))
                X[c] = (B[c] + L.E[k] * X[c + ld] + L.W[k] * X[c - ld] + L.N[k] * X[c + 1] + L.S[k] * X[c - 1]) *
                       L.iP[k];
        }
        if (s < 3) tail_sync(cells);
    }
}

template <class T, int C>
__device__ inline void tail_residual(const TailCells<T, C>& L, const T* X, const T* B, T* R, int ld, bool weighted)
{
#pragma unroll
    for (int k = 0; k < C; k++) {
        const int c = L.c[k];
        if (c < 0) continue;
        T r = T(0);
        if (L.fl[k] & 1u) {
            r = B[c] - L.P[k] * X[c] + L.E[k] * X[c + ld] + L.W[k] * X[c - ld] + L.N[k] * X[c + 1] +
                L.S[k] * X[c - 1];
            if (weighted) r = r * L.wq[k];
        }
        R[L.idx[k]] = r;
    }
}

// coarse cell of level l+1 owned by this thread sums its children's residuals
template <class T>
__device__ inline void tail_restrict(const TailCells<T, 1>& Lc, const TailArgs<T>& A, int l, const T* R, T* B)
{
    if (Lc.c[0] < 0) return;
    const int nj = A.nj[l], fi = A.cx[l] ? 2 : 1, fj = A.cy[l] ? 2 : 1;
    const int cnj = A.nj[l + 1];
    const int I = 1 + Lc.idx[0] / cnj, J = 1 + Lc.idx[0] % cnj;
    T sum = T(0);
    for (int a = 0; a < fi; a++)
        for (int b = 0; b < fj; b++) {
            const int i = fi * (I - 1) + 1 + a, j = fj * (J - 1) + 1 + b;
            if (i <= A.ni[l] && j <= nj) sum += R[(i - 1) * nj + (j - 1)];
        }
    B[Lc.c[0]] = sum;
}

template <class T, int C>
__device__ inline void tail_prolong(const TailCells<T, C>& L, T* X)
{
#pragma unroll
    for (int k = 0; k < C; k++)
        if (L.c[k] >= 0 && (L.fl[k] & 1u)) X[L.c[k]] = X[L.c[k]] + X[L.par[k]];
}

// LDS of k_tail_reg: X and B (ncell each), R (rn), then, 8-byte aligned, the
// FP64 coarsest inverse (CINV_LD x MG_COARSEST_MAX) and its gathered b
template <class T>
__host__ __device__ inline size_t tail_ci_offset(int ncell, int rn)
{
    return (((size_t)ncell * 2 + rn) * sizeof(T) + 7) & ~(size_t)7;
}
template <class T>
inline size_t tail_lds_bytes(int ncell, int top)
{
    return tail_ci_offset<T>(ncell, std::max(top, TAILR_NT)) +
           ((size_t)CINV_LD * MG_COARSEST_MAX + MG_COARSEST_MAX) * sizeof(double) + 16;
}

template <class T, int C0>
__global__ void __launch_bounds__(TAILR_NT) k_tail_reg(TailArgs<T> A)
{
    extern __shared__ unsigned char tail_lds[];
    T* X = reinterpret_cast<T*>(tail_lds);
    T* B = X + A.ncell;
    T* R = B + A.ncell;  // residuals of one level, by cell number (>= max(cells0, TAILR_NT) entries)
    const int q = threadIdx.x;
    const int nt = A.nt;
    const int rn = max(A.ni[0] * A.nj[0], TAILR_NT);
    // the coarsest level's FP64 inverse (row pitch CINV_LD) and its gathered
    // right-hand side, 8-byte aligned after the residuals
    double* CI = reinterpret_cast<double*>(tail_lds + tail_ci_offset<T>(A.ncell, rn));
    double* CB = CI + CINV_LD * MG_COARSEST_MAX;
    static_assert(MG_COARSEST_MAX * MG_COARSEST_MAX <= TAILR_NT, "one inverse entry per thread");
    for (int c = q; c < A.ncell; c += TAILR_NT) {
        X[c] = T(0);
        B[c] = T(0);
    }
    // every global load in flight together: the inverse, the FP32 scale,
    // each level's coefficients and activity, the first level's b
    const int cn2 = A.cn * A.cn;
    const double cinv_q = q < cn2 ? A.cinv[q] : 0.0;
    const double s32v = A.s32 ? *A.s32 : 1.0;
    TailCells<T, C0> L0;
    T b0[C0];
    unsigned a0[C0];
    tail_fetch(L0, A, 0, q, b0, a0);
    TailCells<T, 1> Lk[TAIL_REG_LEVELS - 1];  // tail levels 1 ..
    T bk[TAIL_REG_LEVELS - 1][1];
    unsigned ak[TAIL_REG_LEVELS - 1][1];
    int ncl[TAIL_REG_LEVELS];
    ncl[0] = A.ni[0] * A.nj[0];
#pragma unroll
    for (int l = 1; l < TAIL_REG_LEVELS; l++) {
        ncl[l] = l < nt ? A.ni[l] * A.nj[l] : 0;
        if (l < nt) {
            tail_fetch(Lk[l - 1], A, l, q, bk[l - 1], ak[l - 1]);
        } else {
            Lk[l - 1].c[0] = -1;
            ak[l - 1][0] = 0;
        }
    }
    __syncthreads();  // the zeroing is complete before b lands in B
    if (q < cn2) CI[(q / A.cn) * CINV_LD + q % A.cn] = cinv_q;
    const double inv_s32 = 1.0 / s32v;  // a power of two: exact
    tail_finish(L0, 0, b0, a0, B);
#pragma unroll
    for (int l = 1; l < TAIL_REG_LEVELS; l++) tail_finish(Lk[l - 1], l, bk[l - 1], ak[l - 1], B);
    __syncthreads();
    // ---- down
    if (nt > 1) {
        const int ld = A.nj[0] + 2;
        tail_sweeps4(L0, X, B, ld, ncl[0]);
        tail_sync(ncl[0]);
        tail_residual(L0, X, B, R, ld, A.w != nullptr);
        tail_sync(ncl[0]);
        tail_restrict(Lk[0], A, 0, R, B);
        tail_sync(ncl[0]);
    }
#pragma unroll
    for (int l = 1; l < TAIL_REG_LEVELS - 1; l++) {
        if (l + 1 < nt) {
            const int ld = A.nj[l] + 2;
            tail_sweeps4(Lk[l - 1], X, B, ld, ncl[l]);
            tail_sync(ncl[l]);
            tail_residual(Lk[l - 1], X, B, R, ld, false);
            tail_sync(ncl[l]);
            tail_restrict(Lk[l], A, l, R, B);
            tail_sync(ncl[l]);
        }
    }
    // ---- coarsest level (<= MG_COARSEST_MAX cells, all in wavefront 0):
    // x = A^-1 b in FP64 with the inverse built at setup; thread q < cn
    // gathers b_q, then forms row q
    {
        const int l = nt - 1;
        const int nj = A.nj[l], ld = nj + 2, o = A.off[l], cn = A.cn;
        const int cell = q < cn ? o + (1 + q / nj) * ld + 1 + q % nj : 0;
        if (q < cn) CB[q] = (double)B[cell];
        tail_sync(cn);
        if (q < cn) {
            double s = 0.0;
#pragma unroll 8
            for (int t = 0; t < cn; t++) s += CI[q * CINV_LD + t] * CB[t];
            X[cell] = (T)(s * inv_s32);
        }
        if (nt == 1) __syncthreads();  // the output below reads cells other threads solved
    }
    // ---- up
#pragma unroll
    for (int l = TAIL_REG_LEVELS - 2; l >= 1; l--) {
        if (l + 1 < nt) {
            tail_sync(ncl[l]);
            tail_prolong(Lk[l - 1], X);
            tail_sync(ncl[l]);
            tail_sweeps4(Lk[l - 1], X, B, A.nj[l] + 2, ncl[l]);
        }
    }
    if (nt > 1) {
        tail_sync(ncl[0]);
        tail_prolong(L0, X);
        tail_sync(ncl[0]);
        tail_sweeps4(L0, X, B, A.nj[0] + 2, ncl[0]);
    }
    // ---- first tail level out (each thread its own cells)
#pragma unroll
    for (int k = 0; k < C0; k++) {
        if (L0.c[k] < 0) continue;
        const int nj = A.nj[0];
        const int i = 1 + L0.idx[k] / nj, j = 1 + L0.idx[k] % nj;
        A.x_out[(size_t)i * A.gld[0] + j] = X[L0.c[k]];
    }
}

}  // namespace cathode
