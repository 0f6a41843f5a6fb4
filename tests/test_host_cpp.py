"""C++ host side of the drop-in (02_cathode-08_modulerev_amd/host): CmpMat2
and the S <-> ABCD conversions against the oracle and network identities,
the setup.yaml reader (keys, missing-key and type-error reports), the restart
CSV round trip in the reference layout, and the solverFuncs.hpp drop-in
throwing when no device context can be made (CPU only)."""
import os
import subprocess

import pytest

import host_build


@pytest.fixture(scope="module")
def exe():
    return host_build.build_host_check()


def test_host_headers_cpu(exe, tmp_path):
    yaml = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "setup_reference_layout.yaml")
    r = subprocess.run([exe, "cpu", str(tmp_path), yaml], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    for tag in ("ok abcd", "ok yaml", "ok restart"):
        assert tag in r.stdout


def test_dropin_throws_without_device(exe):
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([exe, "noinit"], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 0 and "ok noinit" in r.stdout, r.stdout + r.stderr
