"""bohrium_amd — an MI355X-native vector engine (`hip`) for the Bohrium lazy
array runtime, with a small bh107-style front end to drive it.

    import bohrium_amd as bh
    a = bh.array(numpy_array); b = (a * 2 + 1).sum(); print(b.asnumpy())

Programs queue bytecode through the bhc C bridge into the component stack
configured by BH_STACK (default: the hip engine, see config.ini).
"""
import numpy as np

from . import _bh_api
from . import opcodes
from .bharray import (BhArray, array, empty, full, zeros, ones, empty_like, zeros_like, ones_like, arange, take, put, flush, ufunc,
                      assign, asnumpy_many, Iterator, get_iterator)
from . import random
from . import user_kernel
from . import blas
from .blas import matmul

_GLOBAL_UFUNCS = ["add", "subtract", "multiply", "true_divide", "divide", "floor_divide", "power", "absolute",
                  "greater", "greater_equal", "less", "less_equal", "equal", "not_equal", "logical_and",
                  "logical_or", "logical_xor", "logical_not", "maximum", "minimum", "bitwise_and", "bitwise_or",
                  "bitwise_xor", "invert", "left_shift", "right_shift", "cos", "sin", "tan", "cosh", "sinh",
                  "tanh", "arcsin", "arccos", "arctan", "arcsinh", "arccosh", "arctanh", "exp", "exp2", "expm1",
                  "log", "log2", "log10", "log1p", "sqrt", "ceil", "trunc", "floor", "rint", "isnan", "isinf",
                  "isfinite", "sign", "real", "imag", "conjugate", "arctan2", "fmod", "remainder", "mod"]
for _n in _GLOBAL_UFUNCS:
    globals()[_n] = ufunc(_n)
abs = ufunc("absolute")


def sum(a, axis=None):
    return a.sum(axis)


def prod(a, axis=None):
    return a.prod(axis)


def cumsum(a, axis=0):
    return a.cumsum(axis)


def cumprod(a, axis=0):
    return a.cumprod(axis)


def do_while(func, niters, *args, **kwargs):
    """Run `func` up to `niters` times as one flushed batch (reference bohrium/loop.py:554-618).

    `func` records one iteration; if it returns a boolean BhArray of one
    element, iteration stops as soon as that value is False."""
    flush()
    cond = func(*args, **kwargs)
    if niters is None:
        if cond is None:
            raise ValueError("do_while without a condition needs niters")
        niters = (1 << 62)
    _bh_api.flush_and_repeat(int(niters), None if cond is None else cond.base.handle)
    return cond


def message(msg):
    """Send a message down the component stack ("info", "statistic", ...)."""
    return _bh_api.message(msg)


def asnumpy(a):
    return a.asnumpy() if isinstance(a, BhArray) else np.asarray(a)
