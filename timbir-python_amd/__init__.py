"""MI355X (gfx950) engine for the ICD reconstruction path of TIMBIR.

The compute path is libxt_hip.so (HIP kernels behind the C-ABI of
include/xt_icd.h); this package holds its ctypes binding (``xt_hip``), the
ctypes mirror of the engine-API structs (``xt_abi``) and the seeded synthetic
scan generator (``synth``).  The directory name carries a hyphen, so it is
loaded through :func:`load` rather than a plain import.
"""
import importlib.util
import os
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
PKG_NAME = "timbir_python_amd"


def load():
    """Import this directory as the package ``timbir_python_amd``."""
    if PKG_NAME in sys.modules:
        return sys.modules[PKG_NAME]
    spec = importlib.util.spec_from_file_location(
        PKG_NAME, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[PKG_NAME] = mod
    spec.loader.exec_module(mod)
    return mod
