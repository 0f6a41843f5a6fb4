"""The C ABI library: built, loadable without a GPU, exports every entry
point include/cathode_hip.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle_lib import ROOT, abi

HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ \*]*?\b(cathode_[A-Za-z0-9_]+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("cathode_init", "cathode_struct5_solve", "cathode_sor_solve", "cathode_lu9_solve",
              "cathode_helm5_solve", "cathode_ndd_step"):
        assert n in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(abi.LIB_PATH), "build with __graft_entry__.build()"
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = sorted(declared_functions() - exported)
    assert not missing, f"declared but not exported: {missing}"


def test_library_loads_and_binds_without_gpu():
    lib = ctypes.CDLL(abi.LIB_PATH)
    assert abi.missing_symbols(lib) == []
    abi.bind(lib)
    assert lib.cathode_abi_version() == 1


def test_python_binding_fails_loudly_when_library_missing(tmp_path):
    saved = abi._LIB
    abi._LIB = None
    try:
        with pytest.raises(RuntimeError):
            abi.load(str(tmp_path / "nope.so"))
    finally:
        abi._LIB = saved


def test_binding_covers_every_declared_function():
    bound = {n for n, _, _ in abi.SIGNATURES}
    assert declared_functions() <= bound, sorted(declared_functions() - bound)


def test_entry_points_refuse_before_init():
    """Without cathode_init every compute entry returns CATHODE_ERR_NOINIT
    before touching the device (so this runs without a GPU)."""
    lib = abi.bind(ctypes.CDLL(abi.LIB_PATH))
    z = np.zeros((5, 5))
    box = np.array([1, 3, 1, 3], dtype=np.int32)
    bp = box.ctypes.data_as(abi.c_int_p)
    d = abi.dptr(z)
    st = abi.Status()
    assert lib.cathode_struct5_solve(d, d, d, d, d, d, bp, 1, 5, 5, 5, 10, 1e-10, 0, d,
                                     ctypes.byref(st)) == abi.CATHODE_ERR_NOINIT
    assert b"cathode_init" in lib.cathode_last_error()
    assert lib.cathode_sor_solve(d, d, d, d, d, d, bp, 1, 5, 5, 5, 1.0, 10, 1e-10, d,
                                 ctypes.byref(st)) == abi.CATHODE_ERR_NOINIT
    assert lib.cathode_lu9_solve(*([d] * 10), bp, 1, 5, 5, 5, d, ctypes.byref(st)) == abi.CATHODE_ERR_NOINIT
    assert lib.cathode_helm5_solve(d, d, d, d, d, d, bp, 1, 5, 5, 5, 10, 1e-10, 0, d,
                                   ctypes.byref(st)) == abi.CATHODE_ERR_NOINIT
    p = abi.EthetaParams()
    assert lib.cathode_etheta_solve(ctypes.byref(p), d, d, d, d, None, ctypes.byref(st)) == abi.CATHODE_ERR_NOINIT


@pytest.mark.parametrize("cname,pyname", [
    ("cathode_status", "Status"), ("cathode_ndd_params", "NddParams"), ("cathode_ndd_fields", "NddFields"),
    ("cathode_ndd_report", "NddReport"), ("cathode_etheta_params", "EthetaParams"),
    ("cathode_ndd_segment", "NddSegment"), ("cathode_ndd_geometry", "NddGeometry"),
    ("cathode_mw_params", "MwParams"), ("cathode_mw_fields", "MwFields"), ("cathode_mw_report", "MwReport"),
    ("cathode_neutral_params", "NeutralParams"), ("cathode_neutral_fields", "NeutralFields"),
    ("cathode_neutral_report", "NeutralReport"), ("cathode_ion_params", "IonParams"),
    ("cathode_ion_fields", "IonFields"), ("cathode_ion_report", "IonReport"),
    ("cathode_rate_tables", "RateTables"), ("cathode_transport_params", "TransportParams"),
    ("cathode_transport_fields", "TransportFields"), ("cathode_mw_bc_list", "MwBcList"), ("cathode_mw_bc", "MwBc"),
    ("cathode_mw_solve_opts", "MwSolveOpts"), ("cathode_mw_solve_fields", "MwSolveFields"),
    ("cathode_mw_solve_report", "MwSolveReport")])
def test_ctypes_struct_layout_matches_header(cname, pyname, tmp_path):
    """Every ctypes mirror has the C struct's size and field offsets (gcc on the header)."""
    py = getattr(abi, pyname)
    fields = [f[0] for f in py._fields_]
    src = tmp_path / "lay.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/cathode_hip.h"',
             'int main(void) {', f'printf("%zu\\n", sizeof({cname}));']
    lines += [f'printf("%zu\\n", offsetof({cname}, {f}));' for f in fields]
    lines += ['return 0; }']
    src.write_text("\n".join(lines))
    exe = tmp_path / "lay"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(py)
    assert out[1:] == [getattr(py, f).offset for f in fields]
