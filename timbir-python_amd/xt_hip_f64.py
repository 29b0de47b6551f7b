"""ctypes binding of libxt_hip_f64.so: the library built with Real_t = double, the
reference's REAL_IS_DOUBLE build (Source_Code_4D_Fast/XT_Constants.h:45-53; the
launcher's --real_is_double).  The structs, the run folder's .bin files and every
host array are double; the GPU keeps float arrays and the float build's system
matrix, so results match the double-arithmetic oracle built on that matrix.

This module is xt_hip itself, loaded a second time with the double structs and
library: the same classes and functions (Engine, Icd, engine_run, the seams),
numpy views of struct arrays in float64.
"""
import importlib.util as _ilu
import os as _os
import sys as _sys

_name = __name__ + "_impl"
_spec = _ilu.spec_from_file_location(_name, _os.path.join(_os.path.dirname(__file__), "xt_hip.py"))
_mod = _ilu.module_from_spec(_spec)
_mod.__package__ = __package__
_mod._XT_REAL_IS_DOUBLE = True
_sys.modules[_name] = _mod
_spec.loader.exec_module(_mod)
globals().update({k: v for k, v in vars(_mod).items() if not k.startswith("__")})
assert REAL_IS_DOUBLE  # noqa: F821 (from the loaded module)
