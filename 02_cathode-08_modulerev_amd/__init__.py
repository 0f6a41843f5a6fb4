"""MI355X implicit-solve library for the axisymmetric NDD cathode model.

The compute path is libcathode_hip.so (csrc/, hand-written gfx950 HIP behind a
C ABI declared in include/cathode_hip.h).  This Python package only binds that
ABI (abi.py), builds it (build.py) and synthesises workloads (workloads.py).
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_ROOT = os.path.dirname(PKG_DIR)
