/*
 * cathode_hip.h -- C ABI of the MI355X (gfx950) implicit-solve library.
 *
 * Every entry point is extern "C" with plain pointers and sizes.  Arrays are
 * row-major (ni+2) x ld doubles indexed [i][j] with i = z (axial, slow) and
 * j = r (radial, contiguous); ld >= nj+2.  This is the layout of the
 * reference's vector<vector<double>>(ni+2, vector<double>(nj+2)) fields
 * (include/arrays.hpp:144-278), flattened.  Complex arrays are interleaved
 * (re, im) pairs of doubles with the same indexing.
 *
 * Matrix convention (include/solverFuncs.hpp:77-82): A x = b with
 * A[P][P] = aP and A[P][nb] = -a_nb for nb in {E(i+1), W(i-1), N(j+1), S(j-1)}.
 * Unknowns exist only on the union of n_bl inclusive boxes
 * boxes[4*b + 0..3] = {ilo, ihi, jlo, jhi}; couplings to cells outside the
 * union are dropped; x is written only on in-box cells.
 *
 * Return codes: 0 = CATHODE_OK, negative = error (message from
 * cathode_last_error()).  Non-convergence is NOT an error (it is silent in the
 * reference, include/solverFuncs.hpp:124-141); it is reported through
 * cathode_status.converged and cathode_last_status().
 */
#ifndef CATHODE_HIP_H
#define CATHODE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CATHODE_ABI_VERSION 1

enum {
    CATHODE_OK = 0,
    CATHODE_ERR_ARG = -1,     /* invalid argument (shape, box, null pointer) */
    CATHODE_ERR_HIP = -2,     /* HIP runtime error */
    CATHODE_ERR_NOINIT = -3,  /* cathode_init() not called */
    CATHODE_ERR_NOMEM = -4,   /* device allocation failed */
    CATHODE_ERR_NUMERIC = -5  /* non-finite value met inside a solve */
};

/* precision modes of the multigrid solves */
enum {
    CATHODE_MODE_FP64 = 0,   /* FP64 everywhere */
    CATHODE_MODE_MIXED = 1   /* FP32 V-cycles, FP64 residual / refinement / norms */
};

typedef struct cathode_status {
    int iterations;       /* outer iterations taken (V-cycles or SOR sweeps) */
    int converged;        /* 1 when the stopping test was met */
    double rel_residual;  /* ||b - A x||_2 / ||b||_2 in FP64 (SOR: last max relative change) */
    double seconds;       /* host wall time of the call */
} cathode_status;

/* ------------------------------------------------------------------ */
/* lifecycle                                                          */
/* ------------------------------------------------------------------ */
int cathode_init(int device);     /* selects the device, creates the library stream */
int cathode_finalize(void);       /* frees every cached device buffer */
int cathode_abi_version(void);
const char* cathode_last_error(void);
void cathode_last_status(cathode_status* st);

/* Device memory owned by the caller, for the *_dev entry points (bytes are
 * zeroed at allocation).  kind: 0 host -> device, 1 device -> host,
 * 2 device -> device; copies are complete when the call returns. */
int cathode_dev_alloc(size_t bytes, void** ptr);
int cathode_dev_free(void* ptr);
int cathode_dev_copy(void* dst, const void* src, size_t bytes, int kind);

/* Global sums inside the fused solver kernels: every workgroup of a launch
 * reads the previous launch's per-workgroup partials, which costs
 * (workgroups x partials); above n partials one workgroup pre-reduces them
 * (fixed order) and the launch reads one value per sum.  n < 0 restores the
 * default (4096: grids above about 1024 x 1024 cells); n = 0 pre-reduces
 * always.  Library-wide, set between solves; results agree with either
 * setting to rounding (the summation order differs).  No counterpart in the
 * reference (HYPRE's reductions are its own). */
int cathode_set_presum_min(int n);

/* ------------------------------------------------------------------ */
/* structured 5-point solves (drop-in for include/solverFuncs.hpp)    */
/* ------------------------------------------------------------------ */

/* Replaces solver_SMG (include/solverFuncs.hpp:26-162): solves A x = b on
 * the box union to ||b-Ax||_2/||b||_2 <= tol or maxITR iterations, from a
 * zero initial guess (the reference never sets xvec, :106-109).  Host
 * arrays; x in/out (only in-box cells are written).  mode = CATHODE_MODE_*. */
int cathode_struct5_solve(const double* aP, const double* aE, const double* aW,
                          const double* aN, const double* aS, const double* b,
                          const int* boxes, int n_bl, int ni2, int nj2, int ld,
                          int maxITR, double tol, int mode,
                          double* x, cathode_status* st);

/* Same solve on device-resident arrays (pointers from hipMalloc, same
 * layout); stream may be NULL (library stream).  Used by callers that keep
 * their fields on the GPU. */
int cathode_struct5_solve_dev(const double* aP, const double* aE, const double* aW,
                              const double* aN, const double* aS, const double* b,
                              const int* boxes, int n_bl, int ni2, int nj2, int ld,
                              int maxITR, double tol, int mode,
                              double* x, void* stream, cathode_status* st);

/* No counterpart in the reference: cathode_struct5_solve_dev for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, semantics and result;
 * it calls no hipMalloc / hipFree / hipHostMalloc.  The multigrid levels,
 * Krylov vectors and reduction slots live in a static 16 MiB device array
 * that each GPU keeps until the process ends (its address is taken once per
 * device by cathode_init).  Limits: at most 64 boxes; grids whose hierarchy
 * needs more than 16 MiB (about 68 000 padded cells, e.g. 256 x 256
 * interior cells) are rejected with CATHODE_ERR_ARG and a message giving
 * the bytes needed -- the reference's 194 x 70 grid needs 3.3 MiB.  The
 * tile shapes are the defaults (no autotuning, which creates events).
 * Every call on one GPU shares the array and keeps the last call's
 * contents; calls that overlap on one GPU race (the library serialises its
 * entry points and each call synchronises its stream before returning). */
int cathode_struct5_solve_dev_noalloc(const double* aP, const double* aE, const double* aW,
                                      const double* aN, const double* aS, const double* b,
                                      const int* boxes, int n_bl, int ni2, int nj2, int ld,
                                      int maxITR, double tol, int mode,
                                      double* x, void* stream, cathode_status* st);

/* Replaces solver_SOR (include/solverFuncs.hpp:169-216): lexicographic
 * point SOR (blocks in order, i then j), x_new = (1-w) x + w (sum a_nb x_nb
 * + b)/(aP + 1e-100); stops after a sweep whose max relative change
 * 2|dx|/|x_new + x_old + 1e-6| <= tol or after the sweep that makes the
 * count exceed maxITR.  x is the initial guess (in/out), as in the
 * reference.  Runs as an anti-diagonal wavefront on the GPU, so the
 * update order -- and the result -- are those of the sequential sweep. */
int cathode_sor_solve(const double* aP, const double* aE, const double* aW,
                      const double* aN, const double* aS, const double* b,
                      const int* boxes, int n_bl, int ni2, int nj2, int ld,
                      double alpha, int maxITR, double tol,
                      double* x, cathode_status* st);

/* Replaces solver_LU_9p (include/solverFuncs.hpp:223-378): 9-point system
 * (the four diagonal neighbours EN=(i+1,j+1), ES=(i+1,j-1), WN=(i-1,j+1),
 * WS=(i-1,j-1) added, same sign convention).  The reference factorizes
 * exactly; this solves to ||b-Ax||/||b|| <= 1e-13 (or maxITR = 200
 * iterations) with a 5-point multigrid preconditioned BiCGStab. */
int cathode_lu9_solve(const double* aP, const double* aE, const double* aW,
                      const double* aN, const double* aS,
                      const double* aEN, const double* aES, const double* aWN,
                      const double* aWS, const double* b,
                      const int* boxes, int n_bl, int ni2, int nj2, int ld,
                      double* x, cathode_status* st);

/* Complex 5-point solve (E_theta Helmholtz system of
 * include/emfieldModule.hpp:2414-2454 folded with its BCs :2713-2773, which
 * the reference hands to Eigen SparseLU :2994-2998).  Interleaved complex128
 * arrays.  mode FP64 = complex128 multigrid; MIXED = complex64 V-cycles with
 * complex128 residual refinement. */
int cathode_helm5_solve(const double* aP, const double* aE, const double* aW,
                        const double* aN, const double* aS, const double* b,
                        const int* boxes, int n_bl, int ni2, int nj2, int ld,
                        int maxITR, double tol, int mode,
                        double* x, cathode_status* st);

/* E_theta microwave field with B = 0, assembled, solved and post-processed
 * on the device: EmfieldModule::solve_Microwave's E_theta rows
 * (include/emfieldModule.hpp:2414-2454), PEC / first-order open folds
 * (:2713-2773), ghost cells (:3213-3264) and J = sigma E, Pabs =
 * 0.5 Re(J E*) (:7733-7743).  Arrays are (ni+2)*(nj+2) cells, ld = nj+2:
 * ne real, J and E interleaved complex128.  Pabs (real, may be NULL) gets
 * the interior values; *P_total (may be NULL) the volume integral
 * sum Pabs 2 pi r dx dr [W].  Ghost corners (and j = 0 on the axis) of E are
 * written as 0. */
typedef struct cathode_etheta_params {
    int ni, nj;
    double dx, dr, rmin;
    double omega;               /* angular frequency [rad/s] */
    double nu;                  /* electron collision frequency [1/s] */
    double epsr;                /* background relative permittivity */
    double masse;               /* electron mass [kg] */
    int pec_z;                  /* left wall (z = xL): 1 = PEC, 0 = first-order open;
                                   the right wall is PEC (emfieldModule.hpp:2729-2737) */
    int pec_rmax;               /* outer wall r = rmax: 1 = PEC, 0 = first-order open */
    int maxITR;
    double tol;
    int mode;                   /* CATHODE_MODE_* */
} cathode_etheta_params;

int cathode_etheta_solve(const cathode_etheta_params* p, const double* ne,
                         const double* J, double* E, double* Pabs,
                         double* P_total, cathode_status* st);

/* ------------------------------------------------------------------ */
/* microwave power profile and port normalisation                     */
/* (EmfieldModule::update_energy_profile, include/emfieldModule.hpp:  */
/*  7630-8105, with mWLoss_at_boundary_xL / rR, :8488-8668)           */
/* ------------------------------------------------------------------ */
#define CATHODE_MW_MAX_BLOCKS 16

/* Inclusive index boxes of the blocks (the reference's gc.i/j_flc_bl,
 * gx.i/j_flx_bl, gr.i/j_flr_bl: cells, x faces, r faces), 4 ints per block:
 * {ilo, ihi, jlo, jhi}.  Every loop of the reference runs over the blocks in
 * order, so a cell or face in two boxes counts twice in the power sums, as
 * there. */
typedef struct cathode_mw_params {
    int ni, nj;
    double dx, dr, rmin;
    double omega;                /* Params::omegam = 2 pi fmw [rad/s] */
    double masse;                /* [kg] */
    int n_bl;                    /* blocks; the last one is the coaxial (dielectric) feed */
    int cell_box[4 * CATHODE_MW_MAX_BLOCKS];
    int xface_box[4 * CATHODE_MW_MAX_BLOCKS];
    int rface_box[4 * CATHODE_MW_MAX_BLOCKS];
    int flag_mwRef;              /* 1: normalise the forward power at the amplifier through the
                                    port's S-parameters; 0: normalise the absorbed power */
    double Pmw;                  /* microwave power [W] */
    double Z0;                   /* line impedance [ohm] */
    double r1, r2;               /* inner / outer radius of the coaxial feed [m] */
    double S11[2], S12[2], S21[2], S22[2];   /* port S-parameters (re, im), Params::S11..S22 */
    double J1r_exc;              /* excitation amplitude, rescaled like the reference's pm.J1r_exc */
    int i_ref;                   /* x face of the reference plane (int(x_ref/dx + 0.5) + 1, x_ref = 0) */
    int j_out, i_out_lo, i_out_hi;  /* cells (i_out_lo..i_out_hi, j_out) under the outer boundary
                                       whose Poynting flux is the loss P_out (gc.j_flc_bl[4][1] - 1,
                                       gc.i_flc_bl[4][0..1]) */
    int j_in_lo, j_in_hi;        /* cells of the reference plane for the Poynting input P_in
                                    (gc.j_flc_bl[n_bl-1][0..1]) */
} cathode_mw_params;

/* Host arrays, each (ni+2)*(nj+2) values, ld = nj+2; complex arrays are
 * interleaved (re, im) doubles.  In: rhoe, nu_m (electron momentum
 * collision frequency nu_m1), Bx, Br (cells) and the field E1x (x faces),
 * E1r (r faces), E1p (cells).  In/out, written inside the blocks' boxes and
 * kept elsewhere as the reference's persistent members: J1x, J1r, J1p =
 * sigma E with the magnetised cold-plasma conductivity tensor, Pabsx, Pabsr,
 * Pabsp = 0.5 Re(J E*) per component, Pabs (cells) = the face averages plus
 * Pabsp.  Rescaled on the whole grid: Pabs (both branches); with
 * flag_mwRef = 1 also E1x, E1r, E1p, J1x/r/p and the excitation currents
 * J1x_exc, J1r_exc, J1p_exc (may be NULL). */
typedef struct cathode_mw_fields {
    const double *rhoe, *nu_m, *Bx, *Br;
    double *E1x, *E1r, *E1p;
    double *J1x, *J1r, *J1p;
    double *J1x_exc, *J1r_exc, *J1p_exc;
    double *Pabsx, *Pabsr, *Pabsp, *Pabs;
} cathode_mw_fields;

typedef struct cathode_mw_report {
    double Px, Pr, Pp;           /* absorbed power per component [W] (before normalisation) */
    double power_sum;            /* absorbed power after normalisation [W] */
    double P_out, P_in;          /* Poynting loss through the outer boundary (after normalisation);
                                    Poynting input at the reference plane (diagnostic, before) [W] */
    double V_ref[2], I_ref[2], Z_ref[2];   /* voltage, current, impedance at the reference plane */
    double P_fwd, P_ref;         /* forward / reflected power at the cathode port [W] */
    double P_fwd_all, P_ref_all; /* the same at the amplifier, through the S-parameters [W] */
    double P_err;                /* P_fwd - P_ref - absorbed - P_out, folded into the absorbed power */
    double ratio, ratio_eff;     /* power ratio Pmw / P_fwd_all (or Pmw / absorbed) and the field scale^2 */
    double Gamma_p[2], Zp[2], Gamma_all[2], Zall[2];  /* reflection coefficients and impedances */
    double J1r_exc;              /* the rescaled excitation amplitude */
} cathode_mw_report;

/* Replaces EmfieldModule::update_energy_profile (emfieldModule.hpp:7630-8105):
 * J = sigma E and the absorbed power on the device, then the normalisation of
 * the flag_mwRef branch (reference-plane V, I and Poynting flux, S-parameter
 * transfer to the amplifier, Pmw / P_fwd_all with the power error folded into
 * the absorbed power) or of the plain branch (Pmw / absorbed power). */
int cathode_mw_energy_profile(const cathode_mw_params* p, cathode_mw_fields* f, cathode_mw_report* rep);

/* ------------------------------------------------------------------ */
/* coupled (E_x, E_r, E_phi) magnetised microwave solve               */
/* (EmfieldModule::solve_Microwave, include/emfieldModule.hpp:2249-3321; */
/*  EmfieldModule::solve_Microwave_impedanceTest, :5528-6141)          */
/* ------------------------------------------------------------------ */

/* Types of the boundary-list entries (MicrowaveBC::sBnd??, the values that
 * makeBoundary writes, include/initialModule.hpp:207-1060).  A type a fold or
 * ghost rule does not name is left alone there, as in the reference. */
#define CATHODE_MWBC_DIRICHLET 0   /* conducting wall (ghost mirror or zero) */
#define CATHODE_MWBC_OPEN 1        /* first-order absorbing boundary */
#define CATHODE_MWBC_GAUSS 2       /* div E = 0 in the boundary cell */
#define CATHODE_MWBC_CONCAVE_A 3   /* concave corner (lower / left wall) */
#define CATHODE_MWBC_CONCAVE_B 4   /* concave corner (upper / right wall) */
#define CATHODE_MWBC_CONVEX_A 5    /* convex corner (lower / left side open) */
#define CATHODE_MWBC_CONVEX_B 6    /* convex corner (upper / right side open) */
#define CATHODE_MWBC_EXCITE 7      /* TEM excitation of the coaxial feed (E_r, west side) */

/* One boundary list (MicrowaveBC iBnd??, jBnd??, sBnd??): entry k is the
 * face or cell (i[k], j[k]) with type s[k]. */
typedef struct cathode_mw_bc_list {
    int n;
    const int* i;
    const int* j;
    const int* s;
} cathode_mw_bc_list;

/* The twelve lists, side index 0 = west (i - 1), 1 = east (i + 1), 2 = south
 * (j - 1), 3 = north (j + 1): x[] for E_x (x faces), r[] for E_r (r faces),
 * p[] for E_phi (cells); MicrowaveBC's iBndWx .. iBndNp.  Folds and ghost
 * rules run in the reference's order: x W, E, S, N, then r, then p. */
typedef struct cathode_mw_bc {
    cathode_mw_bc_list x[4], r[4], p[4];
} cathode_mw_bc;

#define CATHODE_MW_COUPLED 0       /* solve_Microwave: all three components, coupled */
#define CATHODE_MW_IMPEDANCE 1     /* solve_Microwave_impedanceTest: E_r alone (5-point), rhoe = 0,
                                      nu_m = nu_eff, then the reference-plane normalisation */

typedef struct cathode_mw_solve_opts {
    int which;          /* CATHODE_MW_COUPLED or CATHODE_MW_IMPEDANCE */
    int j_flc0;         /* gc.j_flc_bl[0][0]: the r-face type-3 south fold and ghost keep their
                           scale only where r[j] >= j_flc0, the reference's comparison of a
                           radius with this index (emfieldModule.hpp:2673, :3181) */
    double nu_eff;      /* CATHODE_MW_IMPEDANCE: the collision frequency written into nu_m */
    int max_refine;     /* refinement steps after the direct solve, 0..8 */
    double tol;         /* refinement stops at ||b - A E|| / ||b|| <= tol */
} cathode_mw_solve_opts;

/* Arrays (ni+2)*(nj+2), ld = nj+2, complex ones interleaved (re, im).
 * rhoe, nu_m, Bx, Br, epsr: cells (gc.rhoe, nu_m1, Bx, Br, epsr); with
 * CATHODE_MW_IMPEDANCE rhoe and nu_m are overwritten with 0 and nu_eff on the
 * whole grid, as the reference does (:5531-5536).  Er_dc: the static radial
 * field gr.Er (r faces), read only by the east open ghost of E_r (:3162;
 * NULL reads 0).  act_x, act_r, act_p: 1 marks an unknown (jdgBnd_Ex, _Er,
 * _Ep).  J1x_exc, J1r_exc, J1p_exc: excitation currents (complex).  E1x,
 * E1r, E1p (complex, in/out): the unknowns are written, then the ghost values
 * the lists name; every other value is kept (the ghost rules read them).  With
 * CATHODE_MW_IMPEDANCE all three fields and the excitation currents are
 * rescaled on the whole grid by the power ratio. */
typedef struct cathode_mw_solve_fields {
    double *rhoe, *nu_m;
    const double *Bx, *Br, *epsr, *Er_dc;
    const int *act_x, *act_r, *act_p;
    double *J1x_exc, *J1r_exc, *J1p_exc;
    double *E1x, *E1r, *E1p;
} cathode_mw_solve_fields;

typedef struct cathode_mw_solve_report {
    int n_x, n_r, n_p;           /* unknowns per component */
    int block;                   /* dense block size of the column chain */
    int refine;                  /* refinement steps taken */
    double rel_residual;         /* ||b - A E||_2 / ||b||_2 of the solved field (before scaling) */
    /* CATHODE_MW_IMPEDANCE (solve_Microwave_impedanceTest, :5917-6139) */
    double V_ref[2], I_ref[2], Z_ref[2];   /* reference plane (x face p->i_ref) */
    double P_fwd, P_ref;         /* at the cathode port, after scaling [W] */
    double P_fwd_all, P_ref_all; /* at the amplifier, through the S-parameters, after scaling [W] */
    double P_out;                /* Poynting flux through the east face of block 0 (mWLoss_at_boundary_xR) */
    double P_err;                /* P_fwd - P_ref - P_out */
    double ratio;                /* Pmw / P_fwd before scaling; the fields scale by sqrt(ratio) */
    double J1r_exc;              /* the rescaled excitation amplitude */
} cathode_mw_solve_report;

/* Replaces EmfieldModule::solve_Microwave (emfieldModule.hpp:2249-3321) or,
 * with which = CATHODE_MW_IMPEDANCE, solve_Microwave_impedanceTest
 * (:5528-6141).  Of p it reads ni, nj, dx, dr, rmin, omega, masse, the
 * blocks' boxes (the rows are assembled on the boxes' union; a box is
 * x faces 1..ni+1 x 1..nj, r faces 0..ni+1 x 2..nj, cells 0..ni+1 x 1..nj),
 * J1r_exc, r1, r2 (the excitation); with CATHODE_MW_IMPEDANCE also i_ref,
 * Z0, Pmw and the S-parameters.  The system (Eigen SparseLU in the
 * reference) is solved directly on the device: block cyclic reduction over
 * the x columns with dense blocks, then refinement.  Errors: a list entry
 * outside 1..ni x 1..nj (x 2..nj for E_r), two entries of one list on the
 * same face or cell, or a singular block. */
int cathode_mw_solve(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o,
                     cathode_mw_solve_fields* f, cathode_mw_solve_report* rep);
/* The same with every array of f a device pointer (the lists stay on the host). */
int cathode_mw_solve_dev(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o,
                         cathode_mw_solve_fields* f, cathode_mw_solve_report* rep);

/* ------------------------------------------------------------------ */
/* device-resident implicit NDD step                                  */
/* (FluidModule::solve_phi_couple_wdTe_wSEE_PC, solve_rhoe_wdTe_wSEE_PC,
 *  solve_Te_wdTe_wSEE_PC, include/fluidModule.hpp:2080-4957, with the ion
 *  continuity solved by the same Scharfetter-Gummel scheme at charge +1)  */
/* ------------------------------------------------------------------ */

#define CATHODE_SEE_TABLE_MAX 256

typedef struct cathode_ndd_params {
    int ni, nj;                 /* interior cells in z and r */
    double dx, dr, dt;          /* Params::dx, dr, dt */
    double rmin;                /* Params::rmin (0 puts the axis at the lower face) */
    double V_bias;              /* potential of the right (z = zmax) electrode */
    double mu_e, mu_i;          /* isotropic mobilities (B = 0) [m^2/Vs] */
    double Ti;                  /* ion temperature [K] */
    double masse, massi;        /* [kg] */
    double coefIISEE_ts, coefMISEE_ts, Te_emitSEE_eV, ratioEngy_EISEE_rd;
    double dTe_SEE;             /* EISEE table spacing [K] (SeeVec, inputModule.hpp:651-652) */
    int n_see;                  /* EISEE table length, 0 = no electron-induced emission */
    int energy_on;              /* 1 = solve the electron energy equation */
    int maxITR_HES_phi, maxITR_HES_rhoe, maxITR_HES_rhoeps, maxITR_HES_rhoi;
    double error_cnv_HES_phi, error_cnv_HES_rhoe, error_cnv_HES_rhoeps, error_cnv_HES_rhoi;
    int flag_error;             /* calcRes mode 0..3 (miniFuncs.hpp:292-375) */
    /* linear solves, per equation (solver_SMG's maxITR / error_cnv; setup.yaml
     * solver_setting maxITR_SOR_* / error_cnv_SOR_*) */
    int maxITR_SOR_phi, maxITR_SOR_rhoe, maxITR_SOR_rhoeps, maxITR_SOR_rhoi;
    double error_cnv_SOR_phi, error_cnv_SOR_rhoe, error_cnv_SOR_rhoeps, error_cnv_SOR_rhoi;
    int mode;                   /* CATHODE_MODE_* for every linear solve */
    int learned;                /* 1 = learned initial guess for the first Poisson
                                   correction solve (network from cathode_ndd_set_learned) */
    double see_eb[CATHODE_SEE_TABLE_MAX];
    double see_rd[CATHODE_SEE_TABLE_MAX];
    double see_ts[CATHODE_SEE_TABLE_MAX];
    /* time-step control, FluidModule::check_CFL (fluidModule.hpp:57-115),
     * evaluated before every step as main.cpp:174 does: the dielectric
     * relaxation bound dt_min = eps0 / (e max(n_e) max(mu_e)) over every cell.
     * flag_adp_dt = 1 (setup.yaml simulation_control.flag_adp_dt): each step
     * takes dt = CFL * dt_min; 0: dt stays and the report gives CFL = dt / dt_min. */
    int flag_adp_dt;
    double CFL;
} cathode_ndd_params;

/* Host field set, each (ni+2)*(nj+2) doubles, ld = nj+2.  Face arrays use
 * the reference's GridInterfaceX/R convention: x-face i lies between cells
 * i-1 and i, r-face j between j-1 and j.  Source arrays may be NULL (zero). */
typedef struct cathode_ndd_fields {
    double *phi, *rhoe, *rhoi, *Te, *rhoeps, *rhoe_old;   /* cell centred */
    double *Ex, *rhoUex, *nUex, *Gx, *rhoUix;              /* x faces */
    double *Er, *rhoUer, *nUer, *Gr, *rhoUir;              /* r faces */
    double *rate_ionize, *Pabs, *rate_eloss;               /* sources */
} cathode_ndd_fields;

typedef struct cathode_ndd_report {
    int steps;                  /* time steps taken by the call */
    int solves;                 /* linear solves performed */
    int total_iterations;       /* sum of V-cycle iterations over the solves */
    int max_iterations;         /* largest iteration count of one solve */
    int nonconverged;           /* solves that hit maxITR */
    double max_rel_residual;    /* largest final relative residual */
    /* step-level calcRes norms of the last step (Params::error_*) */
    double error_rhoi, error_phi, error_nUex, error_nUer, error_rhoe, error_rhoUex,
           error_rhoUer, error_rhoeps, error_Gx, error_Gr;
    double I_anode, Ii_anode, Ie_anode;   /* right-electrode current [A] (fluidModule.hpp:3942-3965) */
    double seconds;             /* host wall time of the call */
    /* check_CFL before the last step (Params::dt, CFL): the step's dt, the CFL
     * number, dt_min and the maxima it is built from -- max n_e and max Te
     * over every cell, max |Gamma_e| / n_e over the interior plus the thermal
     * speed at max Te (computed as the reference does; it sets nothing) */
    double dt, CFL, dt_min, rhoe_max, Te_max, ue_max;
    /* iterations summed over the call's steps per solve of a step:
     * [2 eq + pass], eq 0 electrons, 1 ions, 2 energy, 3 Poisson; pass 0 the
     * first PC pass, 1 the later ones */
    int iterations_by_solve[8];
    /* the V-cycle cut each solve of the last step ran with, same order (-1:
     * every level; l >= 0: the V-cycles stopped at level l, smoothed there
     * with no coarser correction -- a mass-dominated operator, judged from
     * the same solve one step earlier).  The oracle leaves it 0. */
    int cut_by_solve[8];
} cathode_ndd_report;

int cathode_ndd_create(const cathode_ndd_params* p, void** handle);

/* ------------------------------------------------------------------ */
/* block geometry of the plasma region (include/initialModule.hpp:239-399 */
/* and the boundary loops of the PC routines in fluidModule.hpp)       */
/* ------------------------------------------------------------------ */
/* A handle starts on the rectangle 1..ni x 1..nj with the reference's
 * block-4 boundary set (left wall phi = 0, right electrode phi = V_bias,
 * open top, axis below).  cathode_ndd_set_geometry replaces it by a union of
 * plasma blocks and lists of boundary segments, as the reference's
 * multi-block cathode does:
 *  - cell_box / xface_box / rface_box: gc.i/j_flc_bl, gx.i/j_flx_bl,
 *    gr.i/j_flr_bl of the plasma blocks (the reference's blocks 0 ..
 *    n_bl - 2), inclusive {i0, i1, j0, j1}; the face boxes hold the faces
 *    between two plasma cells, where the Scharfetter-Gummel fluxes, field
 *    updates and calcRes run;
 *  - seg: the boundary faces in the reference's order (z0, z1, z2, z4, z3,
 *    z5, x0, x2, x1, x4, x5, x6, the order of its ghost writes,
 *    fluidModule.hpp:2152-2249); a segment names plasma cells (cell, lo..hi)
 *    and the side of their face; flag bit 0 (CATHODE_SEG_BIAS) puts an x
 *    wall at V_bias, flag bit 1 (CATHODE_SEG_INLET) marks a high-r wall as
 *    the neutral-gas inlet of cathode_ndd_neutral_update (the reference's x1
 *    wall, whose cells at x <= width_neutIn + x6 feed the gas), flag bit 2
 *    (CATHODE_SEG_UIX_BELOW) moves a high-r wall's tangential Uix reflect
 *    of cathode_ndd_ion_update to the row below its cells (that x1 wall);
 *  - ne: the electron-density ghosts set after the electron routine
 *    (fluidModule.hpp:3873-3913): flag 0 writes 0, flag 1 copies the cell;
 *  - te: the Te ghosts copied after the energy routine (:4906-4946).
 * Every face of a plasma cell must be an inner face or lie in exactly one
 * segment of its side; the call rejects a geometry that breaks this,
 * overlapping boxes, or boxes outside 1..ni x 1..nj. */
#define CATHODE_NDD_MAX_BLOCKS 8
#define CATHODE_NDD_MAX_SEGS 24
enum {
    CATHODE_SEG_X_LOW = 0,   /* wall on the low-x face (face i) of cells (i = cell, j = lo..hi) */
    CATHODE_SEG_X_HIGH = 1,  /* wall on the high-x face (face i + 1) */
    CATHODE_SEG_R_LOW = 2,   /* wall on the low-r face (face j) of cells (i = lo..hi, j = cell) */
    CATHODE_SEG_R_HIGH = 3,  /* wall on the high-r face (face j + 1) */
    CATHODE_SEG_R_OPEN = 4,  /* open high-r face: grad phi = 0, electron flux = ion flux */
    CATHODE_SEG_R_AXIS = 5   /* symmetry axis below the cells: no flux, no ghost */
};
#define CATHODE_SEG_BIAS 1
#define CATHODE_SEG_INLET 2
#define CATHODE_SEG_UIX_BELOW 4
typedef struct cathode_ndd_segment {
    int kind, cell, lo, hi, flag;
} cathode_ndd_segment;
typedef struct cathode_ndd_geometry {
    int n_bl;
    int cell_box[4 * CATHODE_NDD_MAX_BLOCKS];
    int xface_box[4 * CATHODE_NDD_MAX_BLOCKS];
    int rface_box[4 * CATHODE_NDD_MAX_BLOCKS];
    int n_seg, n_ne, n_te;
    cathode_ndd_segment seg[CATHODE_NDD_MAX_SEGS];
    cathode_ndd_segment ne[CATHODE_NDD_MAX_SEGS];
    cathode_ndd_segment te[CATHODE_NDD_MAX_SEGS];
} cathode_ndd_geometry;

/* NULL restores the rectangle.  Call before cathode_ndd_set_state; the
 * learned network is not used on a block geometry (the warm starts are). */
int cathode_ndd_set_geometry(void* handle, const cathode_ndd_geometry* g);

/* The reference's five-block cathode: the block boxes of
 * initialModule.hpp:239-343 from xs = {x1, x2, x3, x4, x5, x6, xR} and
 * rs = {r1, r2, r3, r4, r5, rmax} [m] at spacing dx, dr, and the segment
 * lists of the PC routines.  cells6 / xfaces6 / rfaces6 (24 ints each, may be
 * NULL) receive all six block boxes (the five plasma blocks and the
 * dielectric), as the microwave profile takes them.  Host only. */
int cathode_ndd_reference_geometry(const double* xs, const double* rs, double dx, double dr,
                                   cathode_ndd_geometry* g, int* cells6, int* xfaces6, int* rfaces6);

/* ------------------------------------------------------------------ */
/* neutral and metastable densities (FluidModule::update_rhon,          */
/* update_rhom, include/fluidModule.hpp:4963-5411, 5416-5792)           */
/* ------------------------------------------------------------------ */
/* Each is one implicit r-weighted diffusion solve on the handle's plasma
 * blocks (the same 5-point system shape and solver_SMG boundary as the NDD
 * step, zero initial guess), followed by the face-flux update.
 *  - ground-state neutrals n: D = DnN / n_face with DnN = 2 / (3 sigma)
 *    sqrt(k Tn / (pi m_i)), sigma = pi (2 ri)^2, n_face the mean of the two
 *    old cell values; every wall returns the ion and metastable wall fluxes
 *    (Gamma = -Gamma_i,wall - Gamma_m,wall, explicit); on the inlet wall
 *    (CATHODE_SEG_INLET) the cells with x <= x_in also take -fn_In (the flux
 *    update uses x < x_in, as the reference); the open face passes
 *    max(rL / rR Gamma_old, 0) of the face below; source (nu_excStep +
 *    nu_super) n_e - (nu_excMeta + nu_exc / 2 + nu_ionz) n_e; after the flux
 *    update negative n are reset to max(n_old, 0) in every plasma cell.
 *  - metastables m: D = DmN / n_face of the current neutral density;
 *    every wall (the inlet one too) loses 0.25 vth m implicitly, vth =
 *    sqrt(8 k Tn / (pi m_i)); the open face passes rL / rR Gamma_old; source
 *    n_e (nu_excMeta + nu_exc / 2) - n_e (nu_ionzStep + nu_excStep +
 *    nu_super); negative m are reset to max(m_old, 0) in the cells of the
 *    first guard_blocks blocks before the flux update.
 * The reference calls update_rhon every ndt_n steps and update_rhom every
 * ndt_m steps (src/main.cpp:238-247), neutrals first: `which` selects them. */
typedef struct cathode_neutral_params {
    int which;              /* bit 0: neutrals, bit 1: metastables */
    double dt_n, dt_m;      /* dt * ndt_n, dt * ndt_m [s] */
    double Tn, ri, massi;   /* neutral temperature [K], atom radius [m], atom mass [kg] */
    double DmN;             /* metastable diffusion coefficient x neutral density [1/(m s)] */
    double fn_In;           /* inlet particle flux density [1/(m^2 s)] */
    double x_in;            /* width_neutIn + x6 [m] */
    double xL;              /* low-x domain edge: cell x_i = xL + (i - 1/2) dx */
    int guard_blocks;       /* blocks whose negative m are reset (the reference: 2) */
    int maxITR;             /* each solve (maxITR_SOR_phi) */
    double tol;             /* relative residual (error_cnv_SOR_phi) */
    int mode;               /* CATHODE_MODE_FP64 or CATHODE_MODE_MIXED */
    int flag_error;         /* calcRes norm of error_rhon / error_rhom (0..3) */
} cathode_neutral_params;

/* Host arrays of (ni+2) x (nj+2) doubles, [i][j] with ld = nj + 2 (x-face
 * arrays indexed by the face's i, r-face arrays by the face's j).  In/out:
 * rhon, rhom, the fluxes and the wall fluxes of the species updated (the
 * metastable wall fluxes are also read by the neutral update).  In: rhoe, the
 * rate coefficients (NULL = 0) and the ion wall fluxes rhoUix_wall (x faces),
 * rhoUir_wall (r faces).  A species not selected may leave its arrays NULL. */
typedef struct cathode_neutral_fields {
    double *rhon, *rhom;
    const double *rhoe, *nu_excStep, *nu_super, *nu_excMeta, *nu_exc, *nu_ionz, *nu_ionzStep;
    const double *rhoUix_wall, *rhoUir_wall;
    double *rhoUnx, *rhoUnr, *rhoUnx_wall, *rhoUnr_wall;
    double *rhoUmx, *rhoUmr, *rhoUmx_wall, *rhoUmr_wall;
} cathode_neutral_fields;

typedef struct cathode_neutral_report {
    cathode_status solve_n, solve_m;  /* the two linear solves */
    double error_rhon, error_rhom;    /* calcRes of the new against the old density (pm.error_rhon / _rhom) */
    int neg_rhon, neg_rhom;           /* cells reset from a negative value */
} cathode_neutral_report;

/* update_rhon and / or update_rhom on the handle's grid and geometry; the
 * handle's NDD state is not touched.  rep may be NULL. */
int cathode_ndd_neutral_update(void* handle, const cathode_neutral_params* p, cathode_neutral_fields* f,
                               cathode_neutral_report* rep);
/* The same on device arrays of that layout (no PCIe traffic); no reference
 * counterpart, as cathode_ndd_ion_update_dev. */
int cathode_ndd_neutral_update_dev(void* handle, const cathode_neutral_params* p, cathode_neutral_fields* f,
                                   cathode_neutral_report* rep);

/* ------------------------------------------------------------------ */
/* the ion block of the reference's time step (src/main.cpp:194-213):  */
/* FluidModule::solve_Uix_tmp, solve_Uir_tmp, solve_Uip_tmp            */
/* (include/fluidModule.hpp:123-1398), solve_rhoi_constTe (:1405-1755) */
/* and correct_Ui_constTe (:1763-2072)                                 */
/* ------------------------------------------------------------------ */
/* Three implicit upwind momentum solves give the temporary ion velocity:
 * Uix on the inner x faces, Uir on the inner r faces, Uip in the cells
 * (ionization drag S / n_i, ion-neutral collisions nui_m, e E, the Lorentz
 * force of (Bx, Br), the ion pressure k Ti grad n_i / (m_i n_i) and the
 * centrifugal terms).  At a wall the normal velocity is the energy-
 * conserving sign sqrt(U_old^2 + max(2 e dphi / m_i, 0)) of the adjacent
 * unknown (dphi the potential at that face, minus V_bias at a biased wall),
 * the tangential velocities are 0 (Uip: continuity), the open face passes
 * rL / rR max(Uir_old, 0).  The density then solves an upwind continuity
 * with those velocities and a numerical diffusion D = k (Ti + Te_rep) /
 * (m_i (nui_m + nu_star)), nu_star = sqrt(k Te_rep / m_i) / rmax, and
 * the drifting-Maxwellian wall flux (implicit), zero gradient at the open
 * face; the wall fluxes rhoUix_wall / rhoUir_wall are recorded, negative
 * densities reset to max(old, 0).  Finally the velocities and the fluxes
 * rhoUix / rhoUir of the inner faces (and the open face) take the
 * correction -D grad(n_new - n_old).
 * The boundary operations follow the geometry's segments; with the
 * CATHODE_SEG_UIX_BELOW flag a high-r wall's tangential Uix reflect writes
 * the row below its cells instead of the ghost row above, as the reference
 * does for its x1 wall (fluidModule.hpp:520-525).  Every wall needs an
 * inner face next to it (blocks at least two cells across). */
typedef struct cathode_ion_params {
    double dt_i;            /* dt * ndt_i [s] */
    double Ti, massi;       /* ion temperature [K], ion mass [kg] */
    double V_bias;          /* potential of the biased walls [V] */
    double Te_rep_eV;       /* representative electron temperature of the numerical diffusion [eV] */
    double rmax;            /* domain height [m] */
    int maxITR_Ui;          /* the velocity solves (maxITR_SOR_Ui, error_cnv_SOR_Ui) */
    double tol_Ui;
    int maxITR_rhoi;        /* the density solve (maxITR_SOR_rhoi, error_cnv_SOR_rhoi) */
    double tol_rhoi;
    int mode;               /* CATHODE_MODE_FP64 or CATHODE_MODE_MIXED */
} cathode_ion_params;

/* Host arrays of (ni+2) x (nj+2) doubles, [i][j] with ld = nj + 2.  In/out:
 * rhoi, Uix (x faces), Uir (r faces), Uip; out: rhoUix, rhoUir (inner and
 * open faces), rhoUix_wall, rhoUir_wall (wall faces; other entries keep
 * their values).  In: phi, Ex (x faces), Er (r faces), rate_ionize, nui_m,
 * Bx, Br (cells; NULL = 0). */
typedef struct cathode_ion_fields {
    double *rhoi, *Uix, *Uir, *Uip;
    double *rhoUix, *rhoUir, *rhoUix_wall, *rhoUir_wall;
    const double *phi, *Ex, *Er, *rate_ionize, *nui_m, *Bx, *Br;
} cathode_ion_fields;

typedef struct cathode_ion_report {
    cathode_status solve_Uix, solve_Uir, solve_Uip, solve_rhoi;
    int neg_rhoi;                      /* cells reset from a negative density */
} cathode_ion_report;

/* the ion block on the handle's grid and geometry; the handle's NDD state is
 * not touched.  rep may be NULL. */
int cathode_ndd_ion_update(void* handle, const cathode_ion_params* p, cathode_ion_fields* f,
                           cathode_ion_report* rep);
/* The same on device arrays of that layout (cathode_dev_alloc or any HIP
 * allocation on the handle's device), for a caller that keeps the fields
 * resident: nothing crosses PCIe.  No reference counterpart (the reference
 * holds its fields on the host). */
int cathode_ndd_ion_update_dev(void* handle, const cathode_ion_params* p, cathode_ion_fields* f,
                               cathode_ion_report* rep);

/* ------------------------------------------------------------------ */
/* transport coefficients from the rate tables                         */
/* (FluidModule::update_transport_coef, include/fluidModule.hpp:       */
/*  5799-5976; the tables of InputModule::input_BOLSIG_data,           */
/*  include/inputModule.hpp:1407-1654)                                  */
/* ------------------------------------------------------------------ */
/* Electron tables: per neutral (or metastable) density, one entry per
 * mean-energy step of the BOLSIG+ file, looked up by Te [K] with the step
 * dTe = 2/3 of the energy step in K (miniFuncs.hpp func_Te: linear between
 * entries, the last entry beyond the table; func_dTe: the slope, the last
 * interval beyond).  Ion table nui_m by m_i |U_i|^2 with the step dEi [J]
 * (func_Ei: beyond the table the cross-section of the last entry is kept,
 * rate proportional to the speed).  host/tables.hpp reads the files. */
typedef struct cathode_rate_tables {
    int n_e;                                                  /* entries of each electron table */
    const double *nu_elas, *nu_exc, *nu_excReso, *nu_excMeta, *nu_ionz, *nu_ionzStep, *nu_excStep, *nu_super;
    const double *eloss_n, *eloss_m;                          /* energy-loss rates per neutral / metastable [J m^3/s] */
    double dTe;                                               /* [K] */
    int n_i;                                                  /* entries of the ion table */
    const double* nui_m;
    double dEi;                                               /* [J] */
} cathode_rate_tables;

/* copies the tables to the handle's device (host arrays) */
int cathode_ndd_set_rate_tables(void* handle, const cathode_rate_tables* t);

typedef struct cathode_transport_params {
    double masse, massi, Ti;    /* [kg], [kg], [K] */
    double nu_eff;              /* added to the microwave collision frequency nu_m1 [1/s] */
    int flag_Bohm, flag_Sagdeev;
} cathode_transport_params;

/* Arrays of (ni+2) x (nj+2) doubles, [i][j], ld = nj + 2.  In: rhoe, rhoi,
 * rhon, rhom, Te [K], Bx, Br [T], nu_ano (Bohm), Uix (x faces), Uir (r
 * faces), Uip, rhoUex (x faces), rhoUer (r faces); NULL = 0.  In/out:
 * nu_ano_IAT (written on the plasma cells, read everywhere).  Out (every
 * cell of the extended grid; Mache on the plasma cells): the collision
 * frequencies, nu_m, nu_m1, rate_ionize, nui_m, rate_eloss, delossdeps and
 * the mobility tensor mu_para, mu_xx, mu_rr, mu_xr, mu_perp, mu_min.  An
 * output left NULL is not written. */
typedef struct cathode_transport_fields {
    const double *rhoe, *rhoi, *rhon, *rhom, *Te, *Bx, *Br, *nu_ano, *Uix, *Uir, *Uip, *rhoUex, *rhoUer;
    double *nu_ano_IAT;
    double *nu_elas, *nu_exc, *nu_excReso, *nu_excMeta, *nu_ionz, *nu_ionzStep, *nu_excStep, *nu_super;
    double *nu_m, *nu_m1, *rate_ionize, *nui_m, *rate_eloss, *delossdeps;
    double *mu_para, *mu_xx, *mu_rr, *mu_xr, *mu_perp, *mu_min, *Mache;
} cathode_transport_fields;

/* update_transport_coef on the handle's grid and geometry with the tables
 * of cathode_ndd_set_rate_tables (host arrays) */
int cathode_ndd_transport_update(void* handle, const cathode_transport_params* p, cathode_transport_fields* f);
/* The same on device arrays of that layout; no reference counterpart (for
 * callers that keep the fields resident). */
int cathode_ndd_transport_update_dev(void* handle, const cathode_transport_params* p, cathode_transport_fields* f);

int cathode_ndd_set_state(void* handle, const cathode_ndd_fields* host);
int cathode_ndd_get_state(void* handle, cathode_ndd_fields* host);
int cathode_ndd_step(void* handle, int nsteps, cathode_ndd_report* rep);
int cathode_ndd_destroy(void* handle);

/* Independent instances (e.g. the points of a bias sweep, each its own
 * handle and grid size) stepped side by side: nsteps implicit steps of every
 * handle.  Each handle runs on its own stream; the handles are dealt to
 * `workers` host threads largest-first (LPT on cell count), so up to
 * `workers` instances are in flight on the GPU at once.  The first instance
 * of each grid shape takes its first step alone so the tile autotuner sees
 * an idle GPU; the other instances of that shape reuse its choice.  Results
 * per instance are those of cathode_ndd_step on the same handle.  reps
 * (may be NULL) gets one report per handle.  No reference counterpart: the
 * reference runs one bias point per process. */
int cathode_ndd_batch_step(void** handles, int n, int nsteps, int workers,
                           cathode_ndd_report* reps);

/* ------------------------------------------------------------------ */
/* learned initial guess (no reference counterpart)                   */
/* ------------------------------------------------------------------ */
/* A 6-layer 3x3 convolutional network, 32 hidden channels, BF16 weights and
 * activations with FP32 accumulation on the matrix cores, that predicts the
 * solution of the first Poisson correction system of a step from three
 * feature planes: b / rms(b), r / r_top and the previous step's solution of
 * that system / its rms.  The step uses alpha * guess as the initial guess,
 * alpha minimising |b - alpha A guess| (never worse than the zero guess the
 * reference uses; 0 when the guess does not help), and then solves to the
 * same tolerance, so converged results are unchanged.
 * Device arrays are (ni+2)*(nj+2) doubles, pitch nj+2, ghosts 0. */
int cathode_cnn_num_params(void);
int cathode_cnn_create(int seed, void** net);       /* He-initialised weights */
int cathode_cnn_destroy(void* net);
int cathode_cnn_get_params(void* net, float* host); /* FP32 master weights */
int cathode_cnn_set_params(void* net, const float* host);
/* out = network(feat); feat = 3 planes back to back */
int cathode_cnn_forward_dev(void* net, int ni, int nj, const double* feat, double* out);
/* loss = mean over interior cells of (out - target)^2 and its gradient
 * (FP32, parameter layout of cathode_cnn_get_params), no update */
int cathode_cnn_grad_dev(void* net, int ni, int nj, const double* feat,
                         const double* target, double* loss, float* grad_host);
/* one Adam step (lr, beta 0.9 / 0.999, eps 1e-8) on the mean gradient of a
 * batch: feat = batch x 3 planes, target = batch planes; *loss = mean loss */
int cathode_cnn_train_step_dev(void* net, int ni, int nj, int batch, const double* feat,
                               const double* target, float lr, double* loss);
/* Weight gradients on the matrix cores (on = 0, the default: per-band GEMMs
 * of v_mfma_f32_16x16x16_bf16 over the pixels) or as FP32 VALU reductions
 * (on = 1); both sum the per-band partials in the same fixed order. */
int cathode_cnn_set_wgrad_valu(int on);
/* Attach a network to an NDD handle (NULL detaches); used when
 * params.learned = 1.  The network must outlive its use by the handle. */
int cathode_ndd_set_learned(void* handle, void* net);
/* Training capture: each following step writes the features of its first
 * Poisson correction system (3 planes) and the target (solution / rms) into
 * these device buffers; NULL, NULL stops. */
int cathode_ndd_capture(void* handle, double* feat_dev, double* target_dev);

/* Kernel timing of the dominant stencil kernel inside cathode_ndd_step:
 * when enabled, HIP events bracket every launch of the fine-grid
 * red-black smoothing kernel on the step's stream.  Returns the launch
 * count, the summed duration [ms] and the algorithmic bytes of one
 * launch (coefficients + b + x read + x written at storage width). */
int cathode_ndd_profile(void* handle, int enable);
int cathode_ndd_profile_read(void* handle, int* launches, double* total_ms,
                             double* bytes_per_launch);
/* The tile shape the multigrid uses on fine level `level` for the given
 * precision (mixed = 0: FP64 levels, 1: FP32 levels), after the autotuner
 * has run: rows ti x columns tj per workgroup of nt threads.  Levels from
 * the tail on report 0 x 0.  No counterpart in the reference (diagnostics
 * of this implementation's kernels). */
int cathode_ndd_tile(void* handle, int mixed, int level, int* ti, int* tj, int* nt);

#ifdef __cplusplus
}
#endif
#endif /* CATHODE_HIP_H */
