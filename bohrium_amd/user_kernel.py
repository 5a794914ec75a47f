"""User kernels: run hand-written device code on Bohrium arrays.

Mirrors bh107's `user_kernel` module (bridge/bh107/bh107/user_kernel.py:21-92): `execute` flushes the
queued bytecode and hands the kernel source, the operands and a parameter string to the component
stack; the component named by `tag` compiles and runs it.  For the hip engine (tag "hip") the
source defines

    extern "C" __global__ void execute(double *a0, double *a1, ...)

with one pointer per operand (the operand's whole base array, element type as `dtype_to_c99`), and
`param` gives the launch shape like the reference's OpenCL engine
(ve/opencl/engine_opencl.cpp:632-637): {"global_work_size": [X, Y, Z], "local_work_size": [x, y, z]}
with one to three dimensions.  The grid is ceil(global / local) workgroups per dimension, so a kernel
whose global size is not a multiple of the local size checks its index against the array size.
The compiler command is not used by the hip engine (hiprtc compiles for the current device).
"""
import collections.abc

import numpy as np

from . import _bh_api
from .bharray import BhArray


def execute(kernel_source, operand_list, compiler_command=None, tag="hip", param=None, only_behaving_operands=True):
    """Compile and run `execute()` from `kernel_source` with the arrays of `operand_list` as arguments.

    Raises RuntimeError with the component's message when compiling or launching fails."""
    ops = []
    for op in operand_list:
        if not isinstance(op, BhArray):
            raise TypeError("All operands in `operand_list` must be BhArrays")
        if only_behaving_operands and not isbehaving(op):
            raise TypeError("Operand is not behaving set `only_behaving_operands=False` or use `make_behaving()`")
        ops.append(op.view_desc())

    def parse_param():
        for key, value in (param or {}).items():
            if isinstance(value, collections.abc.Iterable) and not isinstance(value, str):
                value = " ".join(str(v) for v in value)
            yield "%s: %s" % (key, value)

    _bh_api.flush()
    msg = _bh_api.user_kernel(kernel_source, ops, compiler_command or "", tag, "; ".join(parse_param()))
    if len(msg) > 0:
        raise RuntimeError(msg)


def isbehaving(ary):
    """C-contiguous, starting at the first element of its base, and as long as the base."""
    return ary.iscontiguous() and ary.offset == 0 and ary.size == ary.base.nelem


def make_behaving(ary, dtype=None):
    """`ary` itself when it is behaving, else a behaving copy; converted to `dtype` when given."""
    ret = ary if isbehaving(ary) else ary.flatten(always_copy=True)
    if dtype is not None and np.dtype(dtype) != ret.dtype:
        ret = ret.astype(dtype)
    return ret


_C99 = {np.dtype(np.float32): "float", np.dtype(np.float64): "double", np.dtype(np.bool_): "bool"}


def dtype_to_c99(dtype):
    """The C name of `dtype` as the kernel sees it (bridge/bh107/bh107/user_kernel.py:95-107)."""
    dtype = np.dtype(dtype)
    if dtype in _C99:
        return _C99[dtype]
    if dtype.kind in "iu":
        return "%s_t" % dtype.name
    raise TypeError("dtype '%s' unsupported" % str(dtype))


def gen_function_prototype(operand_list, operand_name_list=None):
    """The `execute()` declaration for the arrays of `operand_list` (bridge/bh107/bh107/user_kernel.py:110-120)."""
    names = operand_name_list or ["a%d" % i for i in range(len(operand_list))]
    args = ", ".join("%s *%s" % (dtype_to_c99(a.dtype), n) for a, n in zip(operand_list, names))
    return "#include <stdint.h>\nextern \"C\" __global__ void execute(%s)\n" % args
