"""Import shim: loads the package directory ``heyoka.py_amd/`` as the module ``heyoka_amd``.

The directory name contains a dot, which the import system cannot name directly;
this file registers the package under an importable name.
"""

import importlib.util as _ilu
import os as _os
import sys as _sys

_pkg_dir = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "heyoka.py_amd")
_spec = _ilu.spec_from_file_location(__name__, _os.path.join(_pkg_dir, "__init__.py"),
                                     submodule_search_locations=[_pkg_dir])
_mod = _ilu.module_from_spec(_spec)
_sys.modules[__name__] = _mod
_spec.loader.exec_module(_mod)
