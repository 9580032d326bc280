"""MI355X-native batch/ensemble Taylor integrator with heyoka.py's Python API.

Drop-in names for heyoka.py's batch path (heyoka/__init__.py:78-86, 251-336):
``taylor_adaptive_batch``, ``ensemble_propagate_until/for[_batch]``, the
expression system (``make_vars``, ``par``, ``time``, ``sin``, ...), the batch
events (``t_event_batch``, ``nt_event_batch``, ``event_direction``),
``var_ode_sys`` / ``var_args``, ``taylor_outcome`` and ``model``. The
integration runs in hand-written HIP kernels for gfx950 behind the C-ABI of
include/hyx.h (libhyx.so, loaded through ``_lib``).
"""

from .expr import *  # noqa: F401,F403
from .expr import expression, make_vars, par, time, diff, subs  # noqa: F401
from .decompose import taylor_decompose, order_from_tol  # noqa: F401
from . import model  # noqa: F401
from .var_ode_sys import var_ode_sys, var_args  # noqa: F401
from .cfunc import cfunc  # noqa: F401
from .integrator import taylor_adaptive_batch as _tab, taylor_outcome  # noqa: F401
from .events import event_direction, t_event_batch, nt_event_batch  # noqa: F401
from .c_output import continuous_output_batch, continuous_output_batch_dbl  # noqa: F401
from .ensemble import (  # noqa: F401
    ensemble_propagate_until,
    ensemble_propagate_for,
    ensemble_propagate_grid,
    ensemble_propagate_until_batch,
    ensemble_propagate_for_batch,
    ensemble_propagate_grid_batch,
    ensemble_arrays,
)


def taylor_adaptive_batch(sys, state, **kwargs):
    """Factory with heyoka.py's signature (``heyoka/__init__.py:78-86``); FP64 only."""
    import numpy as _np

    fp_type = kwargs.pop("fp_type", float)
    if fp_type not in (float, _np.float64):
        raise TypeError(f'The floating-point type "{fp_type}" is not recognized/supported')
    return _tab(sys, state, **kwargs)


def recommended_simd_size(fp_type=float):
    """Instances per wavefront of the GPU kernels (one instance per lane)."""
    return 64
