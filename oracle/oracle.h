/*
 * oracle.h -- CPU FP64 restatement of the reference's implicit solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this code: it is the checker that the HIP library is compared against
 * (tests/, __graft_entry__.smoke()) and the CPU baseline that bench.py times
 * beside the GPU.  It is re-derived from the formulas of the reference files
 * cited at each function (paths relative to the reference root); no reference
 * source is copied.
 *
 * Parity pins (what fixes each function independently of this restatement):
 *   orc_band_solve / orc_struct5_direct   brute-force dense solve (numpy) on small grids,
 *                                         residual ||b-Ax||/||b|| at roundoff
 *   orc_sor_solve                         hand-computed sweeps on 2x2 / 3x3 systems, fixed
 *                                         point = direct solution
 *   orc_mg_solve                          agrees with orc_struct5_direct to tolerance
 *   orc_bern_*                            B(0)=1, B'(0)=-1/2, B(-z)-B(z)=z, large-|z| limits
 *   orc_ndd_step (Poisson)                exactly linear potential for n_e=n_i, no flux
 *                                         (closed form for any r-weighting); the discrete
 *                                         Gauss law eps0 div E - e dt div Gamma = e(n_i - n_e)
 *                                         - e dt S after the corrected pass (pins the
 *                                         permittivity's dt, mobility, face density and the
 *                                         source term); uniform plasma: Gamma = -mu_e n E
 *   orc_ndd_step (SG continuity)          exponential 1-D steady profile is reproduced
 *                                         exactly; M-matrix signs; global conservation
 *                                         identity (sum of volume divergence = wall flux)
 *   orc_helm_*                            manufactured solution (second-order convergence);
 *                                         open-boundary reflection closed forms; ghost
 *                                         wall conditions; Pabs closed form 0.5 Re(sigma)|E|^2
 *   orc_ndd_assemble (wall faces)         ion, electron and electron-energy wall fluxes are
 *                                         the Maxwellian moments computed by quadrature
 *                                         (drifting ions; electrons over the sheath barrier,
 *                                         2 kT carried per particle); SEE off
 *   orc_ndd_assemble (energy RHS)         J.E source = 3/5 (-e Gamma E) dx dr for uniform
 *                                         fields and flux; 3/5 (Pabs - n_e eloss) dx dr
 *                                         without them
 *   orc_check_cfl                         uniform n_e: dt_min = eps0/(e n mu) exactly; uniform
 *                                         fluxes: ue = |Gamma|/n + vth (closed forms); the
 *                                         adaptive step's dt = CFL dt_min
 *   orc_ndd_step (Te update)              Te0 / 1.001 (the 1e-3 max n_e floor) and the
 *                                         1000 eV cap after a step that barely moves a
 *                                         uniform state
 *   orc_ndd_assemble (SEE)                constant yields: emitted particle / energy wall
 *                                         flux = yield x primary (ion flux, thermal flux of
 *                                         n_e at the step start, thermal energy flux)
 *   orc_mw_energy_profile                 J = sigma E against the 3x3 solve of the cold
 *                                         electron momentum equation m(nu + i w) v = -e(E +
 *                                         v x B); no absorption when nu = 0; B = 0 absorbed
 *                                         power = Drude Re(sigma) |E|^2 / 2 over the exact
 *                                         volumes; TEM coax wave: the reference-plane
 *                                         Poynting input, V, I and Z in closed form
 *                                         (discrete sin / tan factors, -> |V|^2 / 2Z and
 *                                         eta ln(b/a) / 2 pi); smooth fields: both Poynting
 *                                         fluxes converge at second order to the quadrature
 *                                         of the continuous flux (H from Faraday's law); port:
 *                                         P_fwd - P_ref = Re(V conj(I)) / 2, Zp = V / I,
 *                                         Gamma_all = S11 + S12 S21 Gp / (1 - S22 Gp), the
 *                                         amplifier's forward power = Pmw after scaling and
 *                                         absorbed + boundary loss = P_fwd - P_ref
 *                                         (tests/test_oracle_mw.py).  The phi term of the
 *                                         x-face Poynting flux follows the reference's
 *                                         E1p_tmp * conj(H1p_tmp) (emfieldModule.hpp:8518) as
 *                                         written; it is pinned only by the device matching
 *                                         it, every pin above holds E_phi = 0 there
 *   orc_plant                             plants one named mistake; each check above that
 *                                         covers it then fails (tests/test_oracle_pins.py,
 *                                         tests/test_oracle_mw.py)
 *   orc_mw_solve / orc_mw_rows            the reference's impedance-test output (result_deg0/4.csv,
 *                                         six digits) and its recorded amplifier forward power;
 *                                         plasma part of the rows = i omega mu0 dx^2 dr^2 sigma E
 *                                         of orc_mw_energy_profile; B = 0: E_phi = the orc_helm
 *                                         E_theta system; every fold with a ghost rule equals that
 *                                         ghost substituted into the row; Gauss ghosts leave the
 *                                         boundary cell divergence-free; LU residual at roundoff
 *                                         (tests/test_oracle_mwsolve.py).  Parity unpinned: the
 *                                         convex-corner folds (types 5, 6, no ghost rule) and the
 *                                         E_x slot of the concave E_r folds (one row off their own
 *                                         ghost rules in the reference, kept as written)
 *   orc_abcd_*                            S -> ABCD -> S round trip is the identity
 * See tests/test_oracle_*.py for each check.
 */
#ifndef CATHODE_ORACLE_H
#define CATHODE_ORACLE_H

#include "../include/cathode_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- physical constants (include/constants.hpp:5-8), same values ---- */
#define ORC_BOLTZ 1.380649e-23
#define ORC_E0 1.60217662e-19
#define ORC_EPS0 8.8541878128e-12
#define ORC_MU0 1.2566370614e-6
#define ORC_LIM_EXP 1.0e2   /* fluidModule.hpp:3098 */
#define ORC_EPS_VD 1.0e-3   /* fluidModule.hpp:3099 */

/* ---- Scharfetter-Gummel weights (fluidModule.hpp:3239-3243, 3641-3649) ----
 * bp = vd/(1-exp(min(-z,100)))  weight of the upstream (i-1 / j-1) density
 * bm = vd/(exp(min(z,100))-1)   weight of the downstream (i / j) density
 * Canonical form: the differences are taken with expm1 and z == 0 returns
 * the limit D/h.  literal != 0 evaluates the reference's exp()-1 text. */
double orc_bern_bp(double vd, double z, double D_over_h, int literal);
double orc_bern_bm(double vd, double z, double D_over_h, int literal);
void orc_set_literal_bernoulli(int on);

/* Test hook: plant one named mistake in the restatement (0 restores it), so
 * a test can show that the independent check covering that formula fails. */
enum {
    ORC_PLANT_NONE = 0,
    ORC_PLANT_POISSON_DT = 1,         /* permittivity e dt n mu with 2 dt */
    ORC_PLANT_POISSON_MU = 2,         /* permittivity with the ion mobility */
    ORC_PLANT_POISSON_DENSITY = 3,    /* permittivity face density from n_i */
    ORC_PLANT_POISSON_SOURCE = 4,     /* right-hand side without -e dt S */
    ORC_PLANT_POISSON_FLUX_MU = 5,    /* flux correction with the ion mobility */
    ORC_PLANT_SEE_DENSITY = 6,        /* EISEE from n_e instead of n_e at the step start */
    ORC_PLANT_SEE_ION_SIGN = 7,       /* IISEE with the wrong sign */
    ORC_PLANT_SEE_RD_RATIO = 8,       /* energy EISEE without the rediffused energy ratio */
    ORC_PLANT_SEE_EMIT_ENERGY = 9,    /* emitted energy e T_emit instead of 2 e T_emit */
    ORC_PLANT_MW_HALL_SIGN = 10,      /* x-row phi entry of the conductivity with -h_r */
    ORC_PLANT_MW_POYNTING_CONJ = 11,  /* Poynting fluxes with E H instead of E conj(H) */
    ORC_PLANT_MW_PORT_WAVES = 12,     /* forward and reflected waves at the port swapped */
    ORC_PLANT_MW_RFACE_WEIGHT = 13,   /* r-face power volume weighted with the cell radius */
    ORC_PLANT_GEO_FOLD = 14,          /* Poisson: the folds of the high-r walls left out */
    ORC_PLANT_GEO_WALL_COEF = 15,     /* SG: wall faces keep their implicit coefficient */
    ORC_PLANT_NEUT_SIGMA = 16,        /* neutral D with the cross-section pi ri^2 */
    ORC_PLANT_NEUT_INLET_SIGN = 17,   /* neutral inlet feed with the wrong sign in the system */
    ORC_PLANT_META_VTH = 18,          /* metastable wall loss with sqrt(2 k Tn / m) for the mean speed */
    ORC_PLANT_META_RWALL_Q = 19,      /* metastable high-r wall loss without the face weight qR */
    ORC_PLANT_ION_LORENTZ_SIGN = 20,  /* azimuthal ion momentum with the Lorentz force reversed */
    ORC_PLANT_ION_CORR_TE = 21,       /* velocity correction's diffusion without Te_rep */
    ORC_PLANT_TR_DENO = 22,           /* mobility tensor over 2 + W^2 (the unused deno1) */
    ORC_PLANT_TR_IAT_DENSITY = 23,    /* ion plasma frequency of the IAT rate from n_e */
    ORC_PLANT_MWS_RX_WEIGHT = 24,     /* coupled microwave: E_x entries of the r rows weighted 1/2 */
    ORC_PLANT_MWS_EXCITE_SIGN = 25,   /* coupled microwave: excitation source with -i */
    ORC_PLANT_MWS_GAUSS_RADII = 26,   /* coupled microwave: south Gauss fold of E_r with rL / rR */
    ORC_PLANT_MWS_CONCAVE_DENO = 27   /* coupled microwave: west concave fold of E_x (type 3) over 1 + dx/dr qR */
};
void orc_plant(int which);
int orc_planted(void);

/* ---- 5-point solves with the HYPRE box semantics (solverFuncs.hpp:26-162) ---- */
/* Exact direct solve (banded LU, no pivoting: the SG/Poisson matrices are
 * M-matrices).  Returns 0, or -1 on a zero pivot. */
int orc_struct5_direct(const double* aP, const double* aE, const double* aW,
                       const double* aN, const double* aS, const double* b,
                       const int* boxes, int n_bl, int ni2, int nj2, int ld,
                       double* x);
/* complex128 variant with partial pivoting (interleaved re/im arrays) */
int orc_helm5_direct(const double* aP, const double* aE, const double* aW,
                     const double* aN, const double* aS, const double* b,
                     const int* boxes, int n_bl, int ni2, int nj2, int ld,
                     double* x);
/* 9-point exact solve (solverFuncs.hpp:223-378), partial pivoting */
int orc_lu9_direct(const double* aP, const double* aE, const double* aW,
                   const double* aN, const double* aS, const double* aEN,
                   const double* aES, const double* aWN, const double* aWS,
                   const double* b, const int* boxes, int n_bl, int ni2,
                   int nj2, int ld, double* x);
/* Point SOR exactly as solverFuncs.hpp:169-216; returns the sweep count,
 * *err_out = last max relative change. */
int orc_sor_solve(const double* aP, const double* aE, const double* aW,
                  const double* aN, const double* aS, const double* b,
                  const int* boxes, int n_bl, int ni2, int nj2, int ld,
                  double alpha, int maxITR, double tol, double* x,
                  double* err_out);
/* FP64 residual norm ||b-Ax||_2/||b||_2 over the box union */
double orc_rel_residual(const double* aP, const double* aE, const double* aW,
                        const double* aN, const double* aS, const double* b,
                        const int* boxes, int n_bl, int ni2, int nj2, int ld,
                        const double* x);

/* ---- CPU multigrid: the algorithm of the plain-FP64 GPU path, on the host.
 * Used as the CPU baseline and to cross-check the GPU iteration counts.
 * nthreads > 1 uses OpenMP.  Returns iterations; *relres = final residual. */
int orc_mg_solve(const double* aP, const double* aE, const double* aW,
                 const double* aN, const double* aS, const double* b,
                 const int* boxes, int n_bl, int ni2, int nj2, int ld,
                 int maxITR, double tol, double* x, double* relres);
void orc_mg_set_threads(int nthreads);
/* V-cycle cuts of mass-dominated operators on (1, the default) or off */
void orc_mg_set_cut(int on);
/* the cut the last orc_mg_solve ran with (-1: none) */
int orc_mg_last_cut(void);

/* ---- implicit NDD step (fluidModule.hpp:1405-1755 ion (restated as SG at
 * charge +1), :2080-3075 Poisson, :3082-4000 electrons, :4006-4957 energy).
 * solver: 0 = exact direct solves, 1 = CPU multigrid to tol (the baseline). */
int orc_ndd_step(const cathode_ndd_params* p, cathode_ndd_fields* f,
                 int nsteps, int solver, cathode_ndd_report* rep);

/* One assembly of a single equation from the current state, for tests:
 * which = 0 Poisson (first PC iteration), 1 electrons, 2 ions, 3 energy.
 * Fills aP..aS, b ((ni+2)*(nj+2) each, ld = nj+2) after the BC folds, and
 * the predicted face fluxes into fx / fr. */
/* The plasma geometry of the following orc_ndd_step / orc_ndd_assemble calls
 * (blocks, boundary segments and ghost lists, cathode_hip.h); NULL: the
 * rectangle 1..ni x 1..nj with the block-4 boundary set (the default). */
void orc_ndd_set_geometry(const cathode_ndd_geometry* g);
/* the geometry those calls use on an ni x nj grid */
void orc_ndd_geometry(int ni, int nj, cathode_ndd_geometry* out);
/* calcRes over the union of nbox boxes {i0, i1, j0, j1}, block by block */
double orc_calc_res_boxes(const double* vn, const double* vo, int ld, const int* boxes, int nbox, double eps,
                          int mode);

int orc_ndd_assemble(const cathode_ndd_params* p, const cathode_ndd_fields* f,
                     int which, double* aP, double* aE, double* aW, double* aN,
                     double* aS, double* b, double* fx, double* fr);

/* FluidModule::update_rhon / update_rhom (fluidModule.hpp:4963-5792, orc_neutral.c)
 * on the grid of `grid` (ni, nj, dx, dr, rmin) and the geometry set by
 * orc_ndd_set_geometry.  solver: 0 = exact direct solves, 1 = CPU multigrid
 * to p->tol.  Same arguments and outputs as cathode_ndd_neutral_update. */
int orc_neutral_update(const cathode_ndd_params* grid, const cathode_neutral_params* p, cathode_neutral_fields* f,
                       cathode_neutral_report* rep, int solver);

/* the ion block (fluidModule.hpp:123-2072, orc_ion.c), literally over the
 * five plasma blocks of cathode_ndd_reference_geometry set with
 * orc_ndd_set_geometry (returns -1 for another geometry).  solver: 0 =
 * exact direct solves, 1 = CPU multigrid.  Same arguments and outputs as
 * cathode_ndd_ion_update. */
int orc_ion_update(const cathode_ndd_params* grid, const cathode_ion_params* p, cathode_ion_fields* f,
                   cathode_ion_report* rep, int solver);

/* FluidModule::update_transport_coef (fluidModule.hpp:5799-5976) and the
 * table lookups of miniFuncs.hpp:35-91 (orc_transport.c), over the geometry
 * of orc_ndd_set_geometry.  Same arguments and outputs as
 * cathode_ndd_transport_update, the tables as host arrays. */
int orc_transport_update(const cathode_ndd_params* grid, const cathode_transport_params* p,
                         const cathode_rate_tables* T, cathode_transport_fields* f);
double orc_func_Te_tab(const double* v, int n, double Te, double dTe);
double orc_func_dTe_tab(const double* v, int n, double Te, double dTe);
double orc_func_Ei_tab(const double* v, int n, double Ei, double dEi, double massi);

/* EmfieldModule::update_energy_profile (emfieldModule.hpp:7630-8105): J = sigma E,
 * absorbed power, Poynting fluxes and the port normalisation (orc_mw.c) */
int orc_mw_energy_profile(const cathode_mw_params* p, cathode_mw_fields* f, cathode_mw_report* rep);

/* EmfieldModule::solve_Microwave (emfieldModule.hpp:2249-3321) and
 * solve_Microwave_impedanceTest (:5528-6141), orc_mwsolve.c: the same
 * arguments and outputs as cathode_mw_solve (o->max_refine and o->tol unused:
 * the direct solve is banded LU with partial pivoting).  Returns 0, or a
 * negative code for an argument error or a zero pivot. */
int orc_mw_solve(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o,
                 cathode_mw_solve_fields* f, cathode_mw_solve_report* rep);
/* The folded rows of that system, for tests: for component c (0 = E_x,
 * 1 = E_r, 2 = E_phi) the coefficients of slot k (ORC_MWS_*) at face or cell
 * (i, j) are coef[((c * ORC_MWS_NSLOT + k) * (ni+2) + i) * (nj+2) + j]
 * (complex, interleaved); sizes 3 * ORC_MWS_NSLOT * (ni+2)*(nj+2) complex. */
enum { ORC_MWS_P = 0, ORC_MWS_E, ORC_MWS_W, ORC_MWS_N, ORC_MWS_S, ORC_MWS_NE, ORC_MWS_NW, ORC_MWS_SE,
       ORC_MWS_SW, ORC_MWS_EE, ORC_MWS_WW, ORC_MWS_NN, ORC_MWS_SS, ORC_MWS_B, ORC_MWS_NSLOT };
int orc_mw_rows(const cathode_mw_params* p, const cathode_mw_bc* bc, const cathode_mw_solve_opts* o,
                const cathode_mw_solve_fields* f, double* coef);
/* y = A E over the unknowns of that system (rows of inactive faces and cells 0) */
int orc_mw_apply(const cathode_mw_params* p, const cathode_mw_solve_fields* f, int which, const double* coef,
                 const double* Ex, const double* Er, const double* Ep, double* yx, double* yr, double* yp);

/* check_CFL (fluidModule.hpp:57-115): out = {dt_min, max n_e, max mu, max Te,
 * max |Gamma_e|/n_e + vth(max Te)} of the state */
void orc_check_cfl(const cathode_ndd_params* p, const cathode_ndd_fields* f, double* out);

/* calcRes (miniFuncs.hpp:292-375) over an index box, mode 0..3 */
double orc_calc_res(const double* vnew, const double* vold, int ld, int i0,
                    int i1, int j0, int j1, double eps, int mode);

/* EISEE table lerp func_Te (miniFuncs.hpp:35-46) */
double orc_func_Te(const double* table, int n, double Te, double dTe);

/* ---- E_theta microwave system (emfieldModule.hpp:2414-2454, PEC BCs
 * :2713-2773, ghosts :3213-3264) and Pabs (:7739-7741) ---- */
typedef struct orc_helm_params {
    int ni, nj;
    double dx, dr, rmin;
    double omega, nu, epsr;   /* angular frequency, collision frequency, background eps_r */
    double masse;
    int pec_z, pec_rmax;      /* 1 = PEC at both z walls / at r = rmax; 0 = open (first-order ABC) */
} orc_helm_params;
/* Assembles the folded E_theta system (interleaved complex arrays) from
 * n_e (real) and the excitation current J (complex). */
int orc_helm_assemble(const orc_helm_params* p, const double* ne, const double* J,
                      double* aP, double* aE, double* aW, double* aN, double* aS,
                      double* b);
/* Fills ghost cells of E after a solve (PEC mirrors) and computes
 * Pabs = 0.5 Re(J E*) with J = sigma E per cell; returns the volume
 * integral sum Pabs 2 pi r dx dr. */
double orc_helm_post(const orc_helm_params* p, const double* ne, double* E,
                     double* Pabs);

/* ---- 2x2 ABCD / S-parameter algebra (cmpMat2.hpp:8-34,
 * initialModule.hpp:1689-1739); complex numbers as (re, im) pairs ---- */
void orc_s_to_abcd(const double* S /*8: S11,S12,S21,S22*/, double Z0, double* abcd /*8*/);
void orc_abcd_to_s(const double* abcd, double Z0, double* S);

#ifdef __cplusplus
}
#endif
#endif
