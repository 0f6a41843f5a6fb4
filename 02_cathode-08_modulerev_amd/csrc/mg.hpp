// mg.hpp -- geometric multigrid for 5-point cell-centred systems on gfx950.
//
// Algorithm (the host copy in the CPU baseline runs the same steps):
//   weights    fine rows scaled by a radial weight w_j before aggregation
//              (the conservative form of the r-weighted rows; the reference
//              divides every row by r_j, fluidModule.hpp:2634-2641).  The NDD
//              path passes w_j = r_j; the generic solve infers
//              w_{j+1}/w_j = sum_i aN(i,j) / sum_i aS(i,j+1), clipped to [1e-2, 1e2];
//   hierarchy  2x2 aggregation, coarse couplings = halved sums of the fine
//              couplings crossing the coarse face; fine row excess
//              s = aP - sum(a_nb) >= 0 split into a mass part m (minimum
//              excess over the 5-point neighbourhood, summed) and a boundary
//              part s - m (halved); 5-point on every level;
//   smoother   red-black point Gauss-Seidel, V(2,2);
//   transfers  w-weighted sum restriction, piecewise-constant prolongation;
//   coarsest   the first level of <= 32 cells, solved with its dense
//              inverse (built at setup by Gauss-Jordan elimination);
//   cut        a mass-dominated operator may stop its V-cycles at level 0
//              or at level tcut (MGSolver::cut_hint): that level is smoothed
//              with no coarse correction and no coarser launch runs;
//   kernels    one fused down kernel (2 sweeps + residual + restriction) and
//              one fused up kernel (prolongation + 2 sweeps) per fine level,
//              and one workgroup for all levels of <= 512 cells (vcycle.hpp).
//   outer      BiCGStab in FP64 right-preconditioned by one V-cycle per
//              application, from x = 0 (solverFuncs.hpp:106-109), stopping
//              when the true residual ||b - A x||_2 / ||b||_2 <= tol (HYPRE's
//              relative residual) or after maxITR iterations.
//   mixed      the V-cycle runs on FP32 copies of every level (scaled by a
//              power of two so aP ~ 1); residual, update and norms stay FP64.
#pragma once

#include "common.hpp"
#include "vfuse.hpp"

namespace cathode {

// real path: the coarsest level has at most this many cells and is solved
// with its dense inverse (LDS row pitch CINV_LD, no bank conflicts)
constexpr int MG_COARSEST_MAX = 32;
constexpr int CINV_LD = MG_COARSEST_MAX + 1;

constexpr int MG_MAX_LEVELS = 24;

// V-cycle cuts (MGSolver::cut_hint): the excess fraction a level needs for
// the V-cycle to stop there, level 0 and the coarser candidate
constexpr double CUT_DELTA0 = 0.9;
constexpr double CUT_DELTAC = 0.35;

template <class T>
struct Level {
    int ni = 0, nj = 0, ld = 0;
    int cx = 0, cy = 0;  // coarsen to the next level in i / j
    T *aP = nullptr, *aE = nullptr, *aW = nullptr, *aN = nullptr, *aS = nullptr;
    T *b = nullptr, *x = nullptr;
    T* xu = nullptr;  // up kernel output (the level's corrected, post-smoothed x)
};

struct MGProfile {
    bool enabled = false;
    int launches = 0;
    double total_ms = 0.0;
    double bytes_per_launch = 0.0;
    std::vector<hipEvent_t> ev;  // pairs of events, read after the solve
};

// one coarsening step: fine level f -> coarse level c (and its FP32 copy)
struct CoarsenArgs {
    int fni, fnj, fld, fi, fj;
    const double *faP, *faE, *faW, *faN, *faS, *fm;
    const unsigned char* fact;
    const double* w;
    int cni, cnj, cld;
    double *caP, *caE, *caW, *caN, *caS, *cb, *cx, *cm;
    unsigned char* cact;
    float *gP, *gE, *gW, *gN, *gS;  // FP32 copy of the coarse level (mixed)
    double hx, hy, hb;
};

// all coarse levels of the tail, one workgroup, level by level
constexpr int CHAIN_MAX = 12;  // keeps the kernel argument block under 4 KiB
struct CoarsenChain {
    int n;
    CoarsenArgs L[CHAIN_MAX];
};

// the coarsest level, for its dense inverse
struct CoarseInv {
    int ni, nj, ld;
    const double *aP, *aE, *aW, *aN, *aS;
    const unsigned char* act;
    double* inv64;  // n x n, row-major (applied in FP64 by both precisions' tails)
};

class MGSolver {
public:
    MGSolver() = default;
    ~MGSolver();
    MGSolver(const MGSolver&) = delete;
    MGSolver& operator=(const MGSolver&) = delete;

    // Shapes the hierarchy for ni x nj interior cells (no-op if unchanged).
    void configure(int ni, int nj, hipStream_t s);

    // The same inside caller-held device memory [base, base + cap), with
    // no allocation: every device buffer is a piece of that region and the
    // scalars the host reads travel through the host block hbuf of hn
    // doubles (a device mirror in the region is copied into it after each
    // synchronisation).  Buffers are zeroed with asynchronous fills on s.
    // Returns the bytes the shape needs; when that exceeds cap (or the host
    // block is too small: returns SIZE_MAX) nothing is configured.  The
    // graph replay and the tile autotuner (which allocate events) are off.
    size_t configure_in(int ni, int nj, unsigned char* base, size_t cap, double* hbuf, size_t hn, hipStream_t s);

    // Level-0 operator, right-hand side, activity mask and solution, in the
    // solver's (ni+2) x ld0 layout.  Callers fill aP..aS, rhs and act before
    // solve(); inactive cells must have act = 0.
    int ld0() const { return L64[0].ld; }
    double* aP() { return L64[0].aP; }
    double* aE() { return L64[0].aE; }
    double* aW() { return L64[0].aW; }
    double* aN() { return L64[0].aN; }
    double* aS() { return L64[0].aS; }
    double* rhs() { return bsol.p; }
    double* sol() { return xsol.p; }
    // tile shape in use on fine level l (0 x 0 from the tail on)
    void tile_shape(int mixed, int l, int* ti, int* tj, int* nt) const;
    // level-0 activity mask; taking the pointer marks the mask as written,
    // so the next setup mirrors it into the level hierarchy
    unsigned char* act()
    {
        act_dirty = true;
        return act0.p;
    }
    // Radial weights w[0..nj+1] on the device; nullptr = infer from the matrix.
    void set_weights(const double* d_w) { user_w = d_w; }

    // Nine-point outer operator (solver_LU_9p systems).  enable_nine(true)
    // allocates op9(): eight arrays in the level-0 layout, n0() apart, holding
    // the couplings E, W, N, S, EN, ES, WN, WS; the BiCGStab products then use
    // them with aP(), while the V-cycle runs on the level-0 5-point arrays that
    // lump_nine() fills from op9 (each diagonal coupling to an active cell
    // split in halves onto its two adjacent faces, so row sums are kept).
    void enable_nine(bool on);
    double* op9() { return nine.p; }
    size_t n0() const { return level_off[1] - level_off[0]; }
    void lump_nine(hipStream_t s);

    // Solves A x = rhs from x = 0 into sol().  Returns iterations; fills st.
    // With keep_guess set, the solve starts from the x the caller left in
    // sol() (interior cells; ghosts 0) instead; the flag clears itself.
    int solve(int maxITR, double tol, int mode, hipStream_t s, cathode_status* st);
    bool keep_guess = false;
    // Warm start for the next solve (ignored when keep_guess is set or on
    // the nine-point path): the least-squares combination of the given
    // vectors, formed after the setup of that solve.
    void set_warm_start(const WarmSet& ws) { warm = ws; }
    // The number of iterations the next solve is expected to take (e.g. the
    // count of the same solve one time step earlier; 0 = unknown): the
    // iterations before the last expected one run back to back, without a
    // convergence check.  A solve that converges earlier than that keeps
    // iterating up to the first check (a more accurate answer, never a
    // less accurate one); one that needs more continues with checks.
    int expected_its = 0;
    // Where the next solve cuts its V-cycles (the caller's choice, from the
    // cut_next() of the same kind of solve one time step earlier; clears
    // itself): -1 visits every level; a cut the solver offers (level 0, or
    // the second candidate described at tcut) is the last level visited,
    // smoothed from x = 0 with no coarse-grid correction below it, so the
    // V-cycle skips the launches of the coarser levels.  Any other value is
    // ignored.  The converged
    // answer is that of the full V-cycle to the solve's tolerance; only the
    // iteration count can change.
    int cut_hint = -1;
    // After a solve: the cut its operator allows, from the row excess
    // fractions delta_l = min over active cells of (aP - sum a_nb) / aP the
    // setup measured on the candidate levels: level 0 when delta_0 >= 0.9
    // (a mass-dominated system: smoothing alone converges), else the second
    // candidate when its delta >= 0.35, else -1.
    int cut_next() const { return cut_nx; }
    // Hands the solution buffer to the caller in exchange for buf (same
    // size, its own allocation, zero ghosts): no copy.  Kernels already
    // launched with the old sol() pointer still read the solution.
    double* exchange_solution(double* buf)
    {
        double* old = xsol.p;
        xsol.p = buf;
        return old;
    }
    bool act_dirty = true;
    // The caller's promise that the next solve's operator (aP..aS and the
    // activity) is bitwise the one of the last setup, in the same precision,
    // and only the right-hand side changed (the later passes of a
    // predictor-corrector loop): that setup's levels, FP32 copies and
    // coarsest inverse are reused.  Clears itself.
    bool same_operator = false;
    // The caller's kernel already did the level-0 preparation for the next
    // solve (couplings to inactive cells zeroed, row excess into prep_s(),
    // max|aP| and |b|^2 block partials into prep_part() in the 64 x 4 block
    // layout); valid for that one solve, the flag clears itself.
    bool prepared = false;
    double* prep_s() { return L64[0].b; }
    double* prep_part() { return part.p; }
    WarmSet warm;
    DevBuf<double> wpart;  // warm-start partials, warm_nb per dot product
    int warm_nb = 0;

    MGProfile prof;  // dispatch events of the finest-level down kernel

    // A device double the caller writes before solve(); it travels with the
    // solve's first host round trip.  When it is below side_cnv the solve
    // stops after that round trip with aborted = true (NaN never stops it).
    // side_cnv = -inf disables the test.
    double* side_slot() { return scal.p + 15; }
    double side_cnv = -INFINITY;
    bool aborted = false;

    int levels() const { return nl; }

    // Tile configuration (index into VT_CFGS) of each fine level, per
    // precision (0 = FP64, 1 = FP32).  Chosen by autotune() on the first
    // solve of a shape, unless CATHODE_TILE=<index> forces one for all levels
    // or set_autotune(false) keeps the defaults.
    int tile(int prec, int level) const { return tiles[prec][level]; }
    void set_autotune(bool on) { autotune_on = on; }
    // Times every configuration's down + up kernels on each fine level
    // (the current hierarchy of the given precision) and keeps the fastest.
    void autotune(int mixed, hipStream_t s);

private:
    void setup(int mode, hipStream_t s, const WarmSet& ws);
    struct CoarsenArgs coarsen_args(int l, const double* w);
    template <class T>
    void vcycle(Level<T>* L, hipStream_t s, const VFuse& fdown = VFuse(), const VFuse& fup = VFuse(), int cut = -1);
    template <class T>
    int tile_of(int l) const { return tiles[sizeof(T) == 4 ? 1 : 0][l]; }
    template <class T>
    float time_tile(Level<T>* L, int l, int cfg, hipStream_t s);
    void prof_events(hipEvent_t& a, hipEvent_t& b);
    double read_scalar(const double* d, hipStream_t s);
    // The setup's last chain launch (coarse levels of the tail and the
    // coarsest inverse) is left pending: the solve's first restart runs it
    // in the same launch.  flush_chain() launches it alone (before anything
    // else reads the coarse levels).
    bool chain_pending = false;
    int pend_mixed = 0;
    CoarsenChain pend_ch{};
    CoarseInv pend_ci{};
    void flush_chain(hipStream_t s);
    // V-cycle cuts: the candidates are level 0 and tcut (level 2 when level
    // 1 coarsens in both directions and level 2 is a fine level, else level
    // 1; -1 when level 1 is in the tail).  The setup launch leaves per-workgroup
    // minima of delta_0 and delta_tcut in dpart; the restart reduces them
    // into the scalar slots the host reads.
    int tcut = -1;
    int cut_nx = -1;
    DevBuf<double> dpart;
    int dpart_n = 0;    // partials per level the buffer holds
    int setup_dnb = 0;  // workgroups of the last setup launch
    bool dpart_ok = false;  // the last full setup measured the fractions
    int setup_dcl = -1;     // the level of the second fraction (-1: not measured)
    bool cut_valid(int c) const { return !use_nine && lt > 0 && (c == 0 || (c > 0 && c == tcut)); }

    int nl = 0;
    int cfg_ni = -1, cfg_nj = -1;
    Level<double> L64[MG_MAX_LEVELS];
    Level<float> L32[MG_MAX_LEVELS];
    DevBuf<double> store64;   // every FP64 level array
    DevBuf<float> store32;    // every FP32 level array
    DevBuf<double> mass;      // mass part m per level (same layout as the levels)
    DevBuf<unsigned char> actall;  // activity per level
    size_t level_off[MG_MAX_LEVELS + 1] = {0};
    DevBuf<unsigned char> act0;
    DevBuf<double> bsol, xsol, xalt, w0;
    const double* user_w = nullptr;
    DevBuf<double> nine;       // nine-point couplings (enable_nine)
    bool use_nine = false;
    int tiles[2][MG_MAX_LEVELS] = {};
    // captured fused iterations (graph replays of the same launches), keyed
    // by everything their arguments depend on
    struct IterGraph {
        hipGraphExec_t exec = nullptr;
        int mixed = 0, parity = 0, cut = -1;
        const double *xa = nullptr, *pa = nullptr, *w = nullptr;
        int tiles[MG_MAX_LEVELS] = {};
    };
    std::vector<IterGraph> graphs;
    const IterGraph* find_graph(int mixed, int parity, const double* xa, const double* pa, int cut) const;
    template <class Fn>
    const IterGraph* capture_graph(int mixed, int parity, const double* xa, const double* pa, int cut, hipStream_t s,
                                   Fn&& enqueue);
    void drop_graphs();
    bool tuned[2] = {false, false};
    bool autotune_on = true;
    int lt = 0;                // first level handled by the one-workgroup tail kernel
    bool pair_on = true;       // fine levels below the finest run two per launch (k_vpair_*)
    int pair_cfg = 3;          // their tile configuration (VP_CFGS; CATHODE_PAIR_TILE picks another)
    int tail_cells = 0;        // LDS cells of the tail levels
    DevBuf<double> part;       // reduction partials (2 slots)
    // Partial vectors that every workgroup of the next launch sums: above
    // presum_min() partials one workgroup reduces them first (fixed order) into
    // these slots, and the consumers read one value per vector, so their
    // cost stays linear in the grid (below it they sum the partials
    // themselves, which saves a launch on small grids)
    enum { PS_RHO1 = 0, PS_R0V = 1, PS_OM = 2, PS_SETUP = 4, PS_WARM = 6, PS_COUNT = 6 + WARM_NP };
    DevBuf<double> presum;
    const double* presummed(const double* src, int n, int nvec, unsigned max_mask, int slot, int& count,
                            hipStream_t s);
    DevBuf<double> cinv64;    // dense inverse of the coarsest level (FP64, used by both precisions)
    DevBuf<double> kv;         // BiCGStab vectors r, r0, p, v, s, t, ph, sh, p (second buffer)
    DevBuf<double> scal;       // device scalars: [0] |b|^2, [1] |r|^2, [2] scale, [3] f32 matrix scale
    double* h_scal = nullptr;  // pinned mirror of scal (arena mode: the caller's host block)
    size_t h_scal_n = 0;
    bool h_owned = true;       // h_scal came from hipHostMalloc
    int setup_mode = -1, setup_ni = -1, setup_nj = -1;  // precision and shape of the last full setup
    // arena mode (configure_in): buffers are pieces of the caller's region
    Arena* arena = nullptr;    // &region in arena mode; a counting arena while sizing a shape
    Arena region;
    unsigned char* arena_base = nullptr;
    DevBuf<double> dmirror;    // arena mode: what kernels store for the host, copied into h_scal
    template <class T>
    void grab(DevBuf<T>& b, size_t count)
    {
        if (arena) b.borrow(arena->take<T>(count), count);
        else b.alloc(count);
    }
    void build(int ni, int nj);      // shapes, level pointers and device buffers (no fills)
    void zero_buffers(hipStream_t s);
    // where kernels store values for the host, and bringing them there
    double* host_dst() { return arena ? dmirror.p : h_scal; }
    void pull_host(size_t n, hipStream_t s);
};

// Above presum_min() partials, kernels whose every workgroup needs a global sum
// or maximum read it from one value that presum_launch forms first:
// dst[v] = the fixed-order sum (maximum where bit v of max_mask is set) of
// src[v * n + q], q < n, in one workgroup
constexpr int PRESUM_MIN_DEFAULT = 4096;
int presum_min();             // the current threshold (cathode_set_presum_min)
void set_presum_min(int n);
void presum_launch(const double* src, int n, int nvec, unsigned max_mask, double* dst, hipStream_t s);

// FP64 residual-norm helper used by tests of the library itself
double rel_residual_dev(const double* aP, const double* aE, const double* aW, const double* aN,
                        const double* aS, const double* b, const double* x, const unsigned char* act,
                        int ni, int nj, int ld, hipStream_t s);

}  // namespace cathode
