"""ctypes mirror of include/cathode_hip.h and the loader of libcathode_hip.so.

The library is the product: this module only binds its C ABI for Python
callers (tests, bench.py, __graft_entry__).  Loading fails loudly when the
shared object is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "csrc", "libcathode_hip.so")

CATHODE_OK = 0
CATHODE_ERR_ARG = -1
CATHODE_ERR_HIP = -2
CATHODE_ERR_NOINIT = -3
CATHODE_ERR_NOMEM = -4
CATHODE_ERR_NUMERIC = -5
CATHODE_MODE_FP64 = 0
CATHODE_MODE_MIXED = 1
SEE_MAX = 256

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)


class Status(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("seconds", ctypes.c_double)]


class NddParams(ctypes.Structure):
    _fields_ = [
        ("ni", ctypes.c_int), ("nj", ctypes.c_int),
        ("dx", ctypes.c_double), ("dr", ctypes.c_double), ("dt", ctypes.c_double),
        ("rmin", ctypes.c_double), ("V_bias", ctypes.c_double),
        ("mu_e", ctypes.c_double), ("mu_i", ctypes.c_double), ("Ti", ctypes.c_double),
        ("masse", ctypes.c_double), ("massi", ctypes.c_double),
        ("coefIISEE_ts", ctypes.c_double), ("coefMISEE_ts", ctypes.c_double),
        ("Te_emitSEE_eV", ctypes.c_double), ("ratioEngy_EISEE_rd", ctypes.c_double),
        ("dTe_SEE", ctypes.c_double), ("n_see", ctypes.c_int), ("energy_on", ctypes.c_int),
        ("maxITR_HES_phi", ctypes.c_int), ("maxITR_HES_rhoe", ctypes.c_int),
        ("maxITR_HES_rhoeps", ctypes.c_int), ("maxITR_HES_rhoi", ctypes.c_int),
        ("error_cnv_HES_phi", ctypes.c_double), ("error_cnv_HES_rhoe", ctypes.c_double),
        ("error_cnv_HES_rhoeps", ctypes.c_double), ("error_cnv_HES_rhoi", ctypes.c_double),
        ("flag_error", ctypes.c_int), ("maxITR_SOR_phi", ctypes.c_int), ("maxITR_SOR_rhoe", ctypes.c_int),
        ("maxITR_SOR_rhoeps", ctypes.c_int), ("maxITR_SOR_rhoi", ctypes.c_int),
        ("error_cnv_SOR_phi", ctypes.c_double), ("error_cnv_SOR_rhoe", ctypes.c_double),
        ("error_cnv_SOR_rhoeps", ctypes.c_double), ("error_cnv_SOR_rhoi", ctypes.c_double),
        ("mode", ctypes.c_int), ("learned", ctypes.c_int),
        ("see_eb", ctypes.c_double * SEE_MAX), ("see_rd", ctypes.c_double * SEE_MAX),
        ("see_ts", ctypes.c_double * SEE_MAX),
        ("flag_adp_dt", ctypes.c_int), ("CFL", ctypes.c_double),
    ]


FIELD_NAMES = ["phi", "rhoe", "rhoi", "Te", "rhoeps", "rhoe_old",
               "Ex", "rhoUex", "nUex", "Gx", "rhoUix",
               "Er", "rhoUer", "nUer", "Gr", "rhoUir",
               "rate_ionize", "Pabs", "rate_eloss"]
STATE_NAMES = FIELD_NAMES[:16]


class NddFields(ctypes.Structure):
    _fields_ = [(n, c_double_p) for n in FIELD_NAMES]


NDD_MAX_BLOCKS = 8
NDD_MAX_SEGS = 24
SEG_X_LOW, SEG_X_HIGH, SEG_R_LOW, SEG_R_HIGH, SEG_R_OPEN, SEG_R_AXIS = range(6)
SEG_BIAS, SEG_INLET, SEG_UIX_BELOW = 1, 2, 4  # segment flag bits


class NddSegment(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("cell", ctypes.c_int), ("lo", ctypes.c_int), ("hi", ctypes.c_int),
                ("flag", ctypes.c_int)]


class NddGeometry(ctypes.Structure):
    _fields_ = [
        ("n_bl", ctypes.c_int),
        ("cell_box", ctypes.c_int * (4 * NDD_MAX_BLOCKS)), ("xface_box", ctypes.c_int * (4 * NDD_MAX_BLOCKS)),
        ("rface_box", ctypes.c_int * (4 * NDD_MAX_BLOCKS)),
        ("n_seg", ctypes.c_int), ("n_ne", ctypes.c_int), ("n_te", ctypes.c_int),
        ("seg", NddSegment * NDD_MAX_SEGS), ("ne", NddSegment * NDD_MAX_SEGS), ("te", NddSegment * NDD_MAX_SEGS),
    ]


class NddReport(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int), ("solves", ctypes.c_int), ("total_iterations", ctypes.c_int),
        ("max_iterations", ctypes.c_int), ("nonconverged", ctypes.c_int),
        ("max_rel_residual", ctypes.c_double),
        ("error_rhoi", ctypes.c_double), ("error_phi", ctypes.c_double),
        ("error_nUex", ctypes.c_double), ("error_nUer", ctypes.c_double),
        ("error_rhoe", ctypes.c_double), ("error_rhoUex", ctypes.c_double),
        ("error_rhoUer", ctypes.c_double), ("error_rhoeps", ctypes.c_double),
        ("error_Gx", ctypes.c_double), ("error_Gr", ctypes.c_double),
        ("I_anode", ctypes.c_double), ("Ii_anode", ctypes.c_double), ("Ie_anode", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("dt", ctypes.c_double), ("CFL", ctypes.c_double), ("dt_min", ctypes.c_double),
        ("rhoe_max", ctypes.c_double), ("Te_max", ctypes.c_double), ("ue_max", ctypes.c_double),
        ("iterations_by_solve", ctypes.c_int * 8),
        ("cut_by_solve", ctypes.c_int * 8),
    ]


class NeutralParams(ctypes.Structure):
    _fields_ = [
        ("which", ctypes.c_int), ("dt_n", ctypes.c_double), ("dt_m", ctypes.c_double),
        ("Tn", ctypes.c_double), ("ri", ctypes.c_double), ("massi", ctypes.c_double), ("DmN", ctypes.c_double),
        ("fn_In", ctypes.c_double), ("x_in", ctypes.c_double), ("xL", ctypes.c_double),
        ("guard_blocks", ctypes.c_int), ("maxITR", ctypes.c_int), ("tol", ctypes.c_double),
        ("mode", ctypes.c_int), ("flag_error", ctypes.c_int),
    ]


NEUTRAL_FIELD_NAMES = ["rhon", "rhom", "rhoe", "nu_excStep", "nu_super", "nu_excMeta", "nu_exc", "nu_ionz",
                       "nu_ionzStep", "rhoUix_wall", "rhoUir_wall", "rhoUnx", "rhoUnr", "rhoUnx_wall", "rhoUnr_wall",
                       "rhoUmx", "rhoUmr", "rhoUmx_wall", "rhoUmr_wall"]


class NeutralFields(ctypes.Structure):
    _fields_ = [(n, c_double_p) for n in NEUTRAL_FIELD_NAMES]


class NeutralReport(ctypes.Structure):
    _fields_ = [("solve_n", Status), ("solve_m", Status), ("error_rhon", ctypes.c_double),
                ("error_rhom", ctypes.c_double), ("neg_rhon", ctypes.c_int), ("neg_rhom", ctypes.c_int)]


def neutral_fields_struct(arrays):
    """NeutralFields pointing into a dict of numpy arrays (missing keys -> NULL)."""
    f = NeutralFields()
    for n in NEUTRAL_FIELD_NAMES:
        setattr(f, n, dptr(arrays.get(n)))
    return f


class IonParams(ctypes.Structure):
    _fields_ = [
        ("dt_i", ctypes.c_double), ("Ti", ctypes.c_double), ("massi", ctypes.c_double), ("V_bias", ctypes.c_double),
        ("Te_rep_eV", ctypes.c_double), ("rmax", ctypes.c_double), ("maxITR_Ui", ctypes.c_int),
        ("tol_Ui", ctypes.c_double), ("maxITR_rhoi", ctypes.c_int), ("tol_rhoi", ctypes.c_double),
        ("mode", ctypes.c_int),
    ]


ION_FIELD_NAMES = ["rhoi", "Uix", "Uir", "Uip", "rhoUix", "rhoUir", "rhoUix_wall", "rhoUir_wall", "phi", "Ex", "Er",
                   "rate_ionize", "nui_m", "Bx", "Br"]


class IonFields(ctypes.Structure):
    _fields_ = [(n, c_double_p) for n in ION_FIELD_NAMES]


class IonReport(ctypes.Structure):
    _fields_ = [("solve_Uix", Status), ("solve_Uir", Status), ("solve_Uip", Status), ("solve_rhoi", Status),
                ("neg_rhoi", ctypes.c_int)]


def ion_fields_struct(arrays):
    """IonFields pointing into a dict of numpy arrays (missing keys -> NULL)."""
    f = IonFields()
    for n in ION_FIELD_NAMES:
        setattr(f, n, dptr(arrays.get(n)))
    return f


RATE_TABLE_NAMES = ["nu_elas", "nu_exc", "nu_excReso", "nu_excMeta", "nu_ionz", "nu_ionzStep", "nu_excStep",
                    "nu_super", "eloss_n", "eloss_m"]


class RateTables(ctypes.Structure):
    _fields_ = ([("n_e", ctypes.c_int)] + [(n, c_double_p) for n in RATE_TABLE_NAMES] +
                [("dTe", ctypes.c_double), ("n_i", ctypes.c_int), ("nui_m", c_double_p), ("dEi", ctypes.c_double)])


def rate_tables_struct(tabs):
    """RateTables pointing into a dict of numpy tables (keys of RATE_TABLE_NAMES
    plus nui_m, dTe, dEi)"""
    t = RateTables()
    t.n_e = len(tabs["nu_elas"])
    for n in RATE_TABLE_NAMES:
        setattr(t, n, dptr(tabs[n]))
    t.dTe = tabs["dTe"]
    t.n_i = len(tabs["nui_m"])
    t.nui_m = dptr(tabs["nui_m"])
    t.dEi = tabs["dEi"]
    return t


class TransportParams(ctypes.Structure):
    _fields_ = [("masse", ctypes.c_double), ("massi", ctypes.c_double), ("Ti", ctypes.c_double),
                ("nu_eff", ctypes.c_double), ("flag_Bohm", ctypes.c_int), ("flag_Sagdeev", ctypes.c_int)]


TRANSPORT_IN = ["rhoe", "rhoi", "rhon", "rhom", "Te", "Bx", "Br", "nu_ano", "Uix", "Uir", "Uip", "rhoUex", "rhoUer"]
TRANSPORT_OUT = ["nu_elas", "nu_exc", "nu_excReso", "nu_excMeta", "nu_ionz", "nu_ionzStep", "nu_excStep", "nu_super",
                 "nu_m", "nu_m1", "rate_ionize", "nui_m", "rate_eloss", "delossdeps", "mu_para", "mu_xx", "mu_rr",
                 "mu_xr", "mu_perp", "mu_min", "Mache"]
TRANSPORT_FIELD_NAMES = TRANSPORT_IN + ["nu_ano_IAT"] + TRANSPORT_OUT


class TransportFields(ctypes.Structure):
    _fields_ = [(n, c_double_p) for n in TRANSPORT_FIELD_NAMES]


def transport_fields_struct(arrays):
    f = TransportFields()
    for n in TRANSPORT_FIELD_NAMES:
        setattr(f, n, dptr(arrays.get(n)))
    return f


MW_MAX_BLOCKS = 16


class MwParams(ctypes.Structure):
    _fields_ = [
        ("ni", ctypes.c_int), ("nj", ctypes.c_int),
        ("dx", ctypes.c_double), ("dr", ctypes.c_double), ("rmin", ctypes.c_double),
        ("omega", ctypes.c_double), ("masse", ctypes.c_double),
        ("n_bl", ctypes.c_int),
        ("cell_box", ctypes.c_int * (4 * MW_MAX_BLOCKS)), ("xface_box", ctypes.c_int * (4 * MW_MAX_BLOCKS)),
        ("rface_box", ctypes.c_int * (4 * MW_MAX_BLOCKS)),
        ("flag_mwRef", ctypes.c_int), ("Pmw", ctypes.c_double), ("Z0", ctypes.c_double),
        ("r1", ctypes.c_double), ("r2", ctypes.c_double),
        ("S11", ctypes.c_double * 2), ("S12", ctypes.c_double * 2), ("S21", ctypes.c_double * 2),
        ("S22", ctypes.c_double * 2),
        ("J1r_exc", ctypes.c_double),
        ("i_ref", ctypes.c_int), ("j_out", ctypes.c_int), ("i_out_lo", ctypes.c_int), ("i_out_hi", ctypes.c_int),
        ("j_in_lo", ctypes.c_int), ("j_in_hi", ctypes.c_int),
    ]


MW_FIELD_NAMES = ["rhoe", "nu_m", "Bx", "Br", "E1x", "E1r", "E1p", "J1x", "J1r", "J1p",
                  "J1x_exc", "J1r_exc", "J1p_exc", "Pabsx", "Pabsr", "Pabsp", "Pabs"]


class MwFields(ctypes.Structure):
    _fields_ = [(n, c_double_p) for n in MW_FIELD_NAMES]


class MwReport(ctypes.Structure):
    _fields_ = [
        ("Px", ctypes.c_double), ("Pr", ctypes.c_double), ("Pp", ctypes.c_double),
        ("power_sum", ctypes.c_double), ("P_out", ctypes.c_double), ("P_in", ctypes.c_double),
        ("V_ref", ctypes.c_double * 2), ("I_ref", ctypes.c_double * 2), ("Z_ref", ctypes.c_double * 2),
        ("P_fwd", ctypes.c_double), ("P_ref", ctypes.c_double),
        ("P_fwd_all", ctypes.c_double), ("P_ref_all", ctypes.c_double), ("P_err", ctypes.c_double),
        ("ratio", ctypes.c_double), ("ratio_eff", ctypes.c_double),
        ("Gamma_p", ctypes.c_double * 2), ("Zp", ctypes.c_double * 2), ("Gamma_all", ctypes.c_double * 2),
        ("Zall", ctypes.c_double * 2),
        ("J1r_exc", ctypes.c_double),
    ]


def mw_fields_struct(arrays):
    """MwFields pointing into a dict of numpy arrays (complex arrays as
    complex128; missing keys -> NULL)."""
    f = MwFields()
    for n in MW_FIELD_NAMES:
        a = arrays.get(n)
        if a is None:
            setattr(f, n, None)
            continue
        if np.iscomplexobj(a):
            a = a.view(np.float64)
        setattr(f, n, dptr(a))
    return f


MWBC_DIRICHLET, MWBC_OPEN, MWBC_GAUSS, MWBC_CONCAVE_A, MWBC_CONCAVE_B, MWBC_CONVEX_A, MWBC_CONVEX_B, \
    MWBC_EXCITE = range(8)
MW_COUPLED, MW_IMPEDANCE = 0, 1


class MwBcList(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("i", c_int_p), ("j", c_int_p), ("s", c_int_p)]


class MwBc(ctypes.Structure):
    _fields_ = [("x", MwBcList * 4), ("r", MwBcList * 4), ("p", MwBcList * 4)]


class MwSolveOpts(ctypes.Structure):
    _fields_ = [("which", ctypes.c_int), ("j_flc0", ctypes.c_int), ("nu_eff", ctypes.c_double),
                ("max_refine", ctypes.c_int), ("tol", ctypes.c_double)]


MW_SOLVE_FIELD_NAMES = ["rhoe", "nu_m", "Bx", "Br", "epsr", "Er_dc", "act_x", "act_r", "act_p",
                        "J1x_exc", "J1r_exc", "J1p_exc", "E1x", "E1r", "E1p"]
MW_SOLVE_INT_FIELDS = ("act_x", "act_r", "act_p")


class MwSolveFields(ctypes.Structure):
    _fields_ = [(n, c_int_p if n in MW_SOLVE_INT_FIELDS else c_double_p) for n in MW_SOLVE_FIELD_NAMES]


class MwSolveReport(ctypes.Structure):
    _fields_ = [
        ("n_x", ctypes.c_int), ("n_r", ctypes.c_int), ("n_p", ctypes.c_int), ("block", ctypes.c_int),
        ("refine", ctypes.c_int), ("rel_residual", ctypes.c_double),
        ("V_ref", ctypes.c_double * 2), ("I_ref", ctypes.c_double * 2), ("Z_ref", ctypes.c_double * 2),
        ("P_fwd", ctypes.c_double), ("P_ref", ctypes.c_double),
        ("P_fwd_all", ctypes.c_double), ("P_ref_all", ctypes.c_double),
        ("P_out", ctypes.c_double), ("P_err", ctypes.c_double), ("ratio", ctypes.c_double),
        ("J1r_exc", ctypes.c_double),
    ]


def mw_bc_struct(lists):
    """MwBc from {"x": [W, E, S, N], "r": [...], "p": [...]}, each side a list
    of (i, j, s) triples.  Returns (struct, keep) -- keep holds the int arrays
    the struct points into."""
    bc, keep = MwBc(), []
    for comp in ("x", "r", "p"):
        sides = lists.get(comp, [[], [], [], []])
        for k in range(4):
            e = np.ascontiguousarray(np.array(sides[k], dtype=np.int32).reshape(-1, 3).T.copy())
            keep.append(e)
            L = getattr(bc, comp)[k]
            L.n = e.shape[1]
            L.i = e[0].ctypes.data_as(c_int_p)
            L.j = e[1].ctypes.data_as(c_int_p)
            L.s = e[2].ctypes.data_as(c_int_p)
    return bc, keep


def mw_solve_fields_struct(arrays):
    """MwSolveFields pointing into a dict of numpy arrays (complex arrays as
    complex128, masks as int32; missing keys -> NULL)."""
    f = MwSolveFields()
    for n in MW_SOLVE_FIELD_NAMES:
        a = arrays.get(n)
        if a is None:
            setattr(f, n, None)
        elif n in MW_SOLVE_INT_FIELDS:
            assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
            setattr(f, n, a.ctypes.data_as(c_int_p))
        else:
            if np.iscomplexobj(a):
                a = a.view(np.float64)
            setattr(f, n, dptr(a))
    return f


class EthetaParams(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("dx", ctypes.c_double),
                ("dr", ctypes.c_double), ("rmin", ctypes.c_double), ("omega", ctypes.c_double),
                ("nu", ctypes.c_double), ("epsr", ctypes.c_double), ("masse", ctypes.c_double),
                ("pec_z", ctypes.c_int), ("pec_rmax", ctypes.c_int), ("maxITR", ctypes.c_int),
                ("tol", ctypes.c_double), ("mode", ctypes.c_int)]


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(c_double_p)


def fields_struct(state):
    """NddFields pointing into a dict of numpy arrays (missing keys -> NULL)."""
    f = NddFields()
    for n in FIELD_NAMES:
        setattr(f, n, dptr(state.get(n)))
    return f


# exported symbols and their signatures: (name, restype, argtypes)
_D = c_double_p
_SOLVE5 = [_D, _D, _D, _D, _D, _D, c_int_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
SIGNATURES = [
    ("cathode_init", ctypes.c_int, [ctypes.c_int]),
    ("cathode_finalize", ctypes.c_int, []),
    ("cathode_abi_version", ctypes.c_int, []),
    ("cathode_last_error", ctypes.c_char_p, []),
    ("cathode_last_status", None, [ctypes.POINTER(Status)]),
    ("cathode_dev_alloc", ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    ("cathode_dev_free", ctypes.c_int, [ctypes.c_void_p]),
    ("cathode_dev_copy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]),
    ("cathode_struct5_solve", ctypes.c_int,
     _SOLVE5 + [ctypes.c_int, ctypes.c_double, ctypes.c_int, _D, ctypes.POINTER(Status)]),
    ("cathode_struct5_solve_dev", ctypes.c_int,
     _SOLVE5 + [ctypes.c_int, ctypes.c_double, ctypes.c_int, _D, ctypes.c_void_p,
                ctypes.POINTER(Status)]),
    ("cathode_mw_energy_profile", ctypes.c_int,
     [ctypes.POINTER(MwParams), ctypes.POINTER(MwFields), ctypes.POINTER(MwReport)]),
    ("cathode_mw_solve", ctypes.c_int,
     [ctypes.POINTER(MwParams), ctypes.POINTER(MwBc), ctypes.POINTER(MwSolveOpts), ctypes.POINTER(MwSolveFields),
      ctypes.POINTER(MwSolveReport)]),
    ("cathode_mw_solve_dev", ctypes.c_int,
     [ctypes.POINTER(MwParams), ctypes.POINTER(MwBc), ctypes.POINTER(MwSolveOpts), ctypes.POINTER(MwSolveFields),
      ctypes.POINTER(MwSolveReport)]),
    ("cathode_struct5_solve_dev_noalloc", ctypes.c_int,
     _SOLVE5 + [ctypes.c_int, ctypes.c_double, ctypes.c_int, _D, ctypes.c_void_p,
                ctypes.POINTER(Status)]),
    ("cathode_sor_solve", ctypes.c_int,
     _SOLVE5 + [ctypes.c_double, ctypes.c_int, ctypes.c_double, _D, ctypes.POINTER(Status)]),
    ("cathode_lu9_solve", ctypes.c_int,
     [_D] * 10 + [c_int_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D,
                  ctypes.POINTER(Status)]),
    ("cathode_helm5_solve", ctypes.c_int,
     _SOLVE5 + [ctypes.c_int, ctypes.c_double, ctypes.c_int, _D, ctypes.POINTER(Status)]),
    ("cathode_etheta_solve", ctypes.c_int,
     [ctypes.POINTER(EthetaParams), _D, _D, _D, _D, _D, ctypes.POINTER(Status)]),
    ("cathode_ndd_create", ctypes.c_int, [ctypes.POINTER(NddParams), ctypes.POINTER(ctypes.c_void_p)]),
    ("cathode_ndd_set_state", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(NddFields)]),
    ("cathode_ndd_get_state", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(NddFields)]),
    ("cathode_ndd_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(NddReport)]),
    ("cathode_ndd_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("cathode_ndd_batch_step", ctypes.c_int,
     [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(NddReport)]),
    ("cathode_cnn_num_params", ctypes.c_int, []),
    ("cathode_cnn_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("cathode_cnn_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("cathode_cnn_get_params", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("cathode_cnn_set_params", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("cathode_cnn_forward_dev", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_void_p]),
    ("cathode_cnn_grad_dev", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_float)]),
    ("cathode_cnn_train_step_dev", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                                  ctypes.POINTER(ctypes.c_double)]),
    ("cathode_ndd_set_learned", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("cathode_ndd_set_geometry", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(NddGeometry)]),
    ("cathode_cnn_set_wgrad_valu", ctypes.c_int, [ctypes.c_int]),
    ("cathode_set_presum_min", ctypes.c_int, [ctypes.c_int]),
    ("cathode_ndd_reference_geometry", ctypes.c_int,
     [c_double_p, c_double_p, ctypes.c_double, ctypes.c_double, ctypes.POINTER(NddGeometry), c_int_p, c_int_p,
      c_int_p]),
    ("cathode_ndd_capture", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("cathode_ndd_set_rate_tables", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RateTables)]),
    ("cathode_ndd_transport_update", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(TransportParams), ctypes.POINTER(TransportFields)]),
    ("cathode_ndd_transport_update_dev", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(TransportParams), ctypes.POINTER(TransportFields)]),
    ("cathode_ndd_ion_update", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(IonParams), ctypes.POINTER(IonFields), ctypes.POINTER(IonReport)]),
    ("cathode_ndd_ion_update_dev", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(IonParams), ctypes.POINTER(IonFields), ctypes.POINTER(IonReport)]),
    ("cathode_ndd_neutral_update_dev", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(NeutralParams), ctypes.POINTER(NeutralFields),
      ctypes.POINTER(NeutralReport)]),
    ("cathode_ndd_neutral_update", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(NeutralParams), ctypes.POINTER(NeutralFields),
      ctypes.POINTER(NeutralReport)]),
    ("cathode_ndd_profile", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("cathode_ndd_profile_read", ctypes.c_int,
     [ctypes.c_void_p, c_int_p, c_double_p, c_double_p]),
    ("cathode_ndd_tile", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_int_p, c_int_p, c_int_p]),
]


def bind(lib):
    for name, res, args in SIGNATURES:
        if not hasattr(lib, name):
            continue  # reported by missing_symbols(); calling it raises AttributeError
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def missing_symbols(lib):
    """Names of SIGNATURES the library does not export."""
    return [n for n, _, _ in SIGNATURES if not hasattr(lib, n)]


_LIB = None


def load(path=LIB_PATH):
    """Load and bind libcathode_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise RuntimeError(
                f"libcathode_hip.so not built ({path}); run __graft_entry__.build() first")
        _LIB = bind(ctypes.CDLL(path))
    return _LIB


class DeviceArray:
    """float64 device buffer owned through the library (cathode_dev_*), so
    no second GPU runtime is needed in the process."""

    def __init__(self, lib, shape):
        self.lib = lib
        self.shape = tuple(int(s) for s in shape)
        self.size = int(np.prod(self.shape))
        self.ptr = ctypes.c_void_p()
        check(lib.cathode_dev_alloc(max(8, self.size * 8), ctypes.byref(self.ptr)), lib)

    def at(self, k):
        """pointer to the k-th sub-array along the first axis"""
        step = self.size // self.shape[0]
        return ctypes.c_void_p(self.ptr.value + 8 * step * k)

    def upload(self, a, k=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        dst = self.ptr if k is None else self.at(k)
        check(self.lib.cathode_dev_copy(dst, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 0), self.lib)

    def download(self, k=None):
        shape = self.shape if k is None else self.shape[1:]
        out = np.zeros(shape)
        src = self.ptr if k is None else self.at(k)
        check(self.lib.cathode_dev_copy(out.ctypes.data_as(ctypes.c_void_p), src, out.nbytes, 1), self.lib)
        return out

    def copy_from(self, k, other, m):
        """self[k] = other[m] (device to device)"""
        nb = 8 * (self.size // self.shape[0])
        check(self.lib.cathode_dev_copy(self.at(k), other.at(m), nb, 2), self.lib)

    def free(self):
        if self.ptr.value:
            self.lib.cathode_dev_free(self.ptr)
            self.ptr = ctypes.c_void_p()


def device_fields(lib, arrays, names, struct_cls):
    """The present arrays of a dict copied into one device allocation: (a
    struct_cls of device pointers, missing names NULL; the DeviceArray; the
    names in it).  download_fields brings them back."""
    present = [n for n in names if arrays.get(n) is not None]
    shape = arrays[present[0]].shape
    D = DeviceArray(lib, (len(present),) + tuple(shape))
    f = struct_cls()
    for n in names:
        setattr(f, n, None)
    for k, n in enumerate(present):
        D.upload(arrays[n], k)
        setattr(f, n, ctypes.cast(D.at(k), c_double_p))
    return f, D, present


def download_fields(D, present):
    return {n: D.download(k) for k, n in enumerate(present)}


class MwSolveDevice:
    """The arrays of a cathode_mw_solve_dev call in device memory, one
    allocation per array (complex as interleaved float64, masks int32)."""

    def __init__(self, lib, arrays):
        self.lib, self.bufs, self.f = lib, {}, MwSolveFields()
        try:
            for n in MW_SOLVE_FIELD_NAMES:
                v = arrays.get(n)
                if v is None:
                    setattr(self.f, n, None)
                    continue
                raw = np.ascontiguousarray(v.view(np.float64) if np.iscomplexobj(v) else v)
                ptr = ctypes.c_void_p()
                check(lib.cathode_dev_alloc(raw.nbytes, ctypes.byref(ptr)), lib)
                self.bufs[n] = (ptr, v.dtype, v.shape, raw.nbytes)
                check(lib.cathode_dev_copy(ptr, raw.ctypes.data_as(ctypes.c_void_p), raw.nbytes, 0), lib)
                setattr(self.f, n, ctypes.cast(ptr, c_int_p if n in MW_SOLVE_INT_FIELDS else c_double_p))
        except Exception:
            self.free()
            raise

    def download(self):
        out = {}
        for n, (ptr, dt, shape, nb) in self.bufs.items():
            a = np.zeros(shape, dtype=dt)
            check(self.lib.cathode_dev_copy(a.ctypes.data_as(ctypes.c_void_p), ptr, nb, 1), self.lib)
            out[n] = a
        return out

    def free(self):
        for ptr, *_ in self.bufs.values():
            self.lib.cathode_dev_free(ptr)
        self.bufs = {}


def check(rc, lib=None):
    if rc != CATHODE_OK:
        lib = lib or _LIB
        msg = lib.cathode_last_error().decode() if lib is not None else ""
        raise RuntimeError(f"cathode call failed ({rc}): {msg}")
