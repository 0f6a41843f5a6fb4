"""The C ABI library: built, loadable without a GPU, exports every entry
point include/cathode_hip.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

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
    with pytest.raises(RuntimeError):
        abi.load(str(tmp_path / "nope.so")) if abi._LIB is None else abi.load.__wrapped__  # noqa
