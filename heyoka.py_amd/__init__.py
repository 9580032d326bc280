from .expr import *  # noqa
from .decompose import taylor_decompose, order_from_tol  # noqa
from . import model  # noqa
