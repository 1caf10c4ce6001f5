"""MI355X (gfx950) solve path for HAMLSS: Python face of the C ABI.

The directory name is not a Python identifier; load it by path, e.g.
    spec = importlib.util.spec_from_file_location("hsps_amd", ".../hybrid-sparse-pde-solver_amd/__init__.py")
"""
import importlib.util as _ilu
import os as _os

_spec = _ilu.spec_from_file_location("hsps_amd_core", _os.path.join(_os.path.dirname(__file__), "hsps_amd.py"))
_core = _ilu.module_from_spec(_spec)
_spec.loader.exec_module(_core)
globals().update({k: getattr(_core, k) for k in dir(_core) if not k.startswith("__")})
