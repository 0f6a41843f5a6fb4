// ndd_core.hpp -- definitions shared by the implicit NDD step (ndd.hip)
// and the neutral / metastable density updates (neutral.hip): physical
// constants, the kernel-side view of a handle, the geometry policies, the
// row helpers that write a 5-point system into the multigrid's level-0
// arrays, and the handle itself.
#pragma once

#include <cstdint>
#include <initializer_list>
#include <memory>
#include <utility>

#include "common.hpp"
#include "geometry.hpp"
#include "learned.hpp"
#include "mg.hpp"

namespace cathode {

constexpr double K_BOLTZ = 1.380649e-23;   // include/constants.hpp:5
constexpr double K_E0 = 1.60217662e-19;    // :6
constexpr double K_EPS0 = 8.8541878128e-12;  // :7
constexpr double K_LIM_EXP = 1.0e2;        // fluidModule.hpp:3098
constexpr double K_EPS_VD = 1.0e-3;        // fluidModule.hpp:3099
constexpr int NB_X = 64, NB_Y = 4;

enum { SP_ELEC = 0, SP_ION = 1, SP_ENERGY = 2 };

// kernel-side view of the problem (passed by value)
struct NddView {
    int ni, nj, ld, ld0;
    double dx, dr, dt;
    double V_bias, mu_i, Ti, masse, massi;
    double coefIISEE_ts, Te_emitSEE_eV, ratioEngy_EISEE_rd, dTe_SEE;
    int n_see;
    const double* r;  // cell-centre radii r[0..nj+1]
    const double *see_eb, *see_rd, *see_ts;
};

struct Mob {
    double mu_star, inv_par, inv_cross;
};

#define IX(i, j) ((size_t)(i) * (size_t)V.ld + (size_t)(j))
#define IX0(i, j) ((size_t)(i) * (size_t)V.ld0 + (size_t)(j))

#define THREAD_IJ(i0, j0)                                        \
    const int j = (j0) + (int)(blockIdx.x * NB_X + threadIdx.x); \
    const int i = (i0) + (int)(blockIdx.y * NB_Y + threadIdx.y);

// ================================================================== geometry
// Face type codes (geometry.hpp) of x face (i, j) and r face (i, j), the
// plasma cells, the potential ghost rules (-1 or (source << 2) | (Dirichlet
// << 1) | bias, source a field index) and the ghost writes after the
// electron and energy routines.  GeoRect: the rectangle 1..ni x 1..nj with
// the block-4 boundary set (left wall phi = 0, right electrode at V_bias, open
// top, axis below), from the indices.  GeoMask: the tables of a block
// geometry (cathode_ndd_set_geometry), pitch ld.
struct GeoRect {
    __device__ __forceinline__ int xf(const NddView& V, int i, int j) const
    {
        return i == 1 ? GF_WLO : (i == V.ni + 1 ? (GF_WHI | GF_BIAS) : GF_INNER);
    }
    __device__ __forceinline__ int rf(const NddView& V, int i, int j) const
    {
        return j == 1 ? GF_AXIS : (j == V.nj + 1 ? GF_OPEN : GF_INNER);
    }
    // cell (i, j) of the interior 1..ni x 1..nj is plasma
    __device__ __forceinline__ bool act(const NddView&, int, int) const { return true; }
    // any cell of the extended grid
    __device__ __forceinline__ bool act_nb(const NddView& V, int i, int j) const
    {
        return i >= 1 && i <= V.ni && j >= 1 && j <= V.nj;
    }
    __device__ __forceinline__ int ghost(const NddView& V, int i, int j) const
    {
        const bool in_i = i >= 1 && i <= V.ni, in_j = j >= 1 && j <= V.nj;
        if (i == 0 && in_j) return (int)((IX(1, j) << 2) | 2);
        if (i == V.ni + 1 && in_j) return (int)((IX(V.ni, j) << 2) | 3);
        if (j == V.nj + 1 && in_i) return (int)(IX(i, V.nj) << 2);
        return -1;
    }
    __device__ __forceinline__ int te_bits(const NddView& V, int i, int j) const
    {
        return (i == 1 ? (GW_XLO | (GW_XLO << 4)) : 0) | (i == V.ni ? (GW_XHI | (GW_XHI << 4)) : 0) |
               (j == 1 ? (GW_RLO | (GW_RLO << 4)) : 0) | (j == V.nj ? (GW_RHI | (GW_RHI << 4)) : 0);
    }
    __device__ __forceinline__ int ne_bits(const NddView&, int, int) const { return 0; }
    // 1 + the block index of a plasma cell
    __device__ __forceinline__ int blk(const NddView&, int, int) const { return 1; }
};

struct GeoMask {
    const uint8_t *xc, *rc, *ac, *teb, *neb, *bk;
    const int* gs;
    __device__ __forceinline__ int xf(const NddView& V, int i, int j) const { return xc[IX(i, j)]; }
    __device__ __forceinline__ int rf(const NddView& V, int i, int j) const { return rc[IX(i, j)]; }
    __device__ __forceinline__ bool act(const NddView& V, int i, int j) const { return ac[IX(i, j)] != 0; }
    __device__ __forceinline__ bool act_nb(const NddView& V, int i, int j) const { return ac[IX(i, j)] != 0; }
    __device__ __forceinline__ int ghost(const NddView& V, int i, int j) const { return gs[IX(i, j)]; }
    __device__ __forceinline__ int te_bits(const NddView& V, int i, int j) const { return teb[IX(i, j)]; }
    __device__ __forceinline__ int ne_bits(const NddView& V, int i, int j) const { return neb[IX(i, j)]; }
    __device__ __forceinline__ int blk(const NddView& V, int i, int j) const { return bk[IX(i, j)]; }
};

// masks of the transverse-flux interpolation: 0 across a wall (and the axis
// for r faces), as the reference's bLr / bRr / bLx / bRx flags
__device__ __forceinline__ double tmask_r(int code)
{
    const int t = code & GF_TYPE;
    return (double)(t != GF_WLO && t != GF_WHI && t != GF_AXIS);
}
__device__ __forceinline__ double tmask_x(int code)
{
    const int t = code & GF_TYPE;
    return (double)(t != GF_WLO && t != GF_WHI);
}
__device__ __forceinline__ bool is_wall(int t) { return t == GF_WLO || t == GF_WHI; }

// The multigrid's level-0 preparation, done by the assembly kernels: the
// couplings to cells outside the plasma zeroed (HYPRE's off-grid
// convention), the row excess s = max(aP - sum a, 0), and the block partials
// of max|aP| (slot 0) and |b|^2 (slot 1) in the multigrid's 2-D block layout
// (this grid's); rows of non-plasma cells get the identity, as the
// multigrid's own preparation gives them.
// one assembled row: the five coefficients and the right-hand side
struct Row5 {
    double P, E, W, N, S, b;
};

template <class G>
__device__ inline void prep_row(const G& g, const NddView& V, int i, int j, Row5& R, double* s_out, double& amax,
                                double& bb)
{
    R.E = g.act_nb(V, i + 1, j) ? R.E : 0.0;
    R.W = g.act_nb(V, i - 1, j) ? R.W : 0.0;
    R.N = g.act_nb(V, i, j + 1) ? R.N : 0.0;
    R.S = g.act_nb(V, i, j - 1) ? R.S : 0.0;
    const double ex = R.P - (R.E + R.W + R.N + R.S);
    s_out[IX0(i, j)] = ex > 0.0 ? ex : 0.0;
    amax = fabs(R.P);
    bb = R.b * R.b;
}

__device__ inline void store_row(const NddView& V, int i, int j, const Row5& R, double* aP, double* aE, double* aW,
                                 double* aN, double* aS, double* b)
{
    const size_t k = IX0(i, j);
    aP[k] = R.P;
    aE[k] = R.E;
    aW[k] = R.W;
    aN[k] = R.N;
    aS[k] = R.S;
    b[k] = R.b;
}

// a cell outside the plasma: the identity row, b = 0
__device__ inline void store_idle_row(const NddView& V, int i, int j, double* aP, double* aE, double* aW, double* aN,
                                      double* aS, double* b, double* s_out)
{
    const size_t k = IX0(i, j);
    aP[k] = 1.0;
    aE[k] = 0.0;
    aW[k] = 0.0;
    aN[k] = 0.0;
    aS[k] = 0.0;
    b[k] = 0.0;
    if (s_out) s_out[k] = INFINITY;
}

__device__ inline void prep_partials(double amax, double bb, double* part)
{
    __shared__ double lds[8];
    double v[2] = {amax, bb};
    const bool mx[2] = {true, false};
    block_reduce_vec<NB_X * NB_Y, 2>(v, mx, lds);
    const double t1 = v[0], t2 = v[1];
    if (lin_tid() == 0) {
        const int nb = (int)(gridDim.x * gridDim.y), bid = (int)(blockIdx.y * gridDim.x + blockIdx.x);
        part[bid] = t1;
        part[nb + bid] = t2;
    }
}

// calcRes pairs (new, old): cells, inner x faces, inner r faces
struct ResPairs {
    const double *n0, *o0, *n1, *o1, *n2, *o2;
};

// faces of the electrode at V_bias: (x face i, rows j0..j1) per segment
constexpr int ANODE_MAX = 4;
struct AnodeSegs {
    int n;
    int fi[ANODE_MAX], j0[ANODE_MAX], j1[ANODE_MAX];
};

struct NddHandle {
    cathode_ndd_params p{};
    NddView V{};
    Mob me{}, mi{};
    hipStream_t s = nullptr;
    size_t n = 0;
    DevBuf<double> fields;  // 19 state/source arrays
    DevBuf<double> work;    // scratch arrays
    DevBuf<double> radii, see;
    DevBuf<double> part, scal;
    DevBuf<double> cfl_part;  // check_CFL block maxima (3 per block of the extended grid)
    DevBuf<double> sgw;     // SG face weights of one PC pass (SgArgs::wgt)
    double* h_scal = nullptr;
    double* F[19] = {nullptr};
    bool has_src[3] = {false, false, false};
    // work arrays
    double *phi_old, *Ex_old, *Er_old, *tmp_c, *tmp_x, *tmp_r, *dphi, *s_old, *sx_old, *sr_old, *Te_old, *rhoi_old;
    double *tmp2_c, *tmp2_x, *tmp2_r;  // the other set of start-of-pass copies (SG passes alternate)
    MGSolver mg;
    cathode_ndd_report rep{};
    int nb = 0;
    // learned initial guess of the first Poisson correction solve of a step
    void* cnn = nullptr;              // network (cathode_ndd_set_learned), not owned
    std::unique_ptr<CnnScratch, void (*)(CnnScratch*)> cnn_sc{nullptr, cnn_scratch_free};  // this handle's activations
    DevBuf<double> lfeat, lguess, xprev, lpart, lscal;
    bool have_xprev = false;
    // learned mode: every solve starts from its own previous-step solution
    // (slot = equation x PC iteration), scaled by the least-squares alpha
    // the last WARM_HIST solutions of each solve, newest first; one
    // allocation each: they trade places with the solver's solution buffer
    static constexpr int WARM_HIST = 3;
    DevBuf<double> wbuf[8][WARM_HIST];
    int nhist[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int last_its[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // each solve's iteration count one step earlier
    // each solve's V-cycle cut, from its operator one step earlier (-1: every
    // level); CATHODE_CUT=0 keeps every V-cycle whole
    int cut[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    bool cut_on = true;
    double* cap_feat = nullptr;       // capture for training (device, caller-owned)
    double* cap_target = nullptr;
    int guesses = 0, guesses_used = 0;
    // the network is used while it lowers the warm-started residual: after
    // a use that did not (residual ratio >= NET_USEFUL), the next NET_PROBE - 1
    // first Poisson solves start from the warm start alone, then it is tried again
    bool net_useful = true, gate_pending = false;
    int net_wait = 0;
    // block geometry (cathode_ndd_set_geometry); the rectangle when !masked
    bool masked = false;
    GeoMask gm{};
    DevBuf<uint8_t> geo_u8;  // x codes, r codes, plasma cells, Te and n_e ghost bits, block indices
    DevBuf<int> geo_gs;      // potential ghost rules
    AnodeSegs anode{};
    // the geometry in segment form (the rectangle's when none is set) and a
    // counter that changes with every cathode_ndd_set_geometry
    cathode_ndd_geometry geo_src{};
    int geo_version = 0;
    // neutral / metastable update (neutral.hip): its planes, allocated on first use
    DevBuf<double> neut;
    DevBuf<int> neut_cnt;
    // ion block (ion.hip): planes, the per-face operation tables and ghost
    // lists of the current geometry (built for geo_version ion_plan_version)
    DevBuf<double> ionp;
    DevBuf<uint8_t> ion_ops;
    DevBuf<int> ion_ghosts;
    int ion_plan_version = -1;
    int ion_ng[3] = {0, 0, 0};
    DevBuf<int> ion_cnt;
    // transport coefficients (transport.hip): the rate tables (10 electron
    // tables of rate_n[0] entries, then the ion table of rate_n[1]), their
    // steps dTe and dEi, the staging planes and the n_e maximum
    DevBuf<double> rate_tab;
    int rate_n[2] = {0, 0};
    double rate_d[2] = {0.0, 0.0};
    DevBuf<double> trp, tr_scal;
};

// launches kernel template K with the handle's geometry policy as the first
// argument
#define NDD_GEO_LAUNCH(H, K, grid, block, ...)                                                   \
    do {                                                                                          \
        if ((H).masked)                                                                           \
            hipLaunchKernelGGL(K<GeoMask>, grid, block, 0, (H).s, (H).gm, __VA_ARGS__);           \
        else                                                                                      \
            hipLaunchKernelGGL(K<GeoRect>, grid, block, 0, (H).s, GeoRect{}, __VA_ARGS__);        \
    } while (0)

enum {
    F_PHI, F_RHOE, F_RHOI, F_TE, F_RHOEPS, F_RHOE_OLD, F_EX, F_RHOUEX, F_NUEX, F_GX, F_RHOUIX,
    F_ER, F_RHOUER, F_NUER, F_GR, F_RHOUIR, F_S, F_PABS, F_ELOSS
};

static inline dim3 gblk() { return dim3(NB_X, NB_Y); }
static inline dim3 ggrid(int ni, int nj) { return dim3(cdiv(nj, NB_X), cdiv(ni, NB_Y)); }

// dst_q = src_q for each listed pair of whole planes, in one launch
void d2d(NddHandle& H, std::initializer_list<std::pair<double*, const double*>> pairs);
// count doubles from device memory d into the handle's host block (synchronises); returns the first
double read_dev(NddHandle& H, const double* d, int count = 1);
// calcRes of three pairs in norm mode (-1: the handle's flag_error); returns
// the max error (host sync); errors also in scal[0..3]
double calcres3(NddHandle& H, const ResPairs& P, double* err_out, int mode = -1);

}  // namespace cathode
