"""MI355X implementation of TACS's finite-element assembly / block-CSR /
GMRES / adjoint path.  Mirrors the reference's Python modules tacs.TACS,
tacs.elements, tacs.constitutive and tacs.functions for that path; the work
runs in the HIP library tacs_amd/lib/libtacs_hip.so (required)."""
from ._lib import lib, LIB_PATH, TACSError  # noqa: F401
from . import TACS, elements, constitutive, functions, meshes  # noqa: F401

lib()  # fail loudly at import when the native library is missing
# measured launch variants (tacs_amd/tuning_table.json, written by
# tacs_amd/autotune.py): every variant gives the same results
TACS.loadTuningTable()
