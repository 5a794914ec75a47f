"""A lazy N-d array over the bhc bridge, in the style of the reference's bh107
front end (bridge/bh107/bh107/bharray.py:49-572, ufuncs.py:257-400).

Operations queue bytecode; data reaches the host only on request
(`asnumpy()`), which syncs the base and flushes.  When the last Python
reference to a base goes away a BH_FREE is queued, so temporaries created and
freed inside one flush become registers in the fused kernel.
"""
import numbers

import numpy as np

from . import _bh_api as api
from . import opcodes as ops


_DTYPE = type(np.dtype(np.float64))


class BhBase:
    """Owner of one bridge base array; queues BH_FREE when collected."""

    __slots__ = ("handle", "dtype", "nelem", "__weakref__")

    def __init__(self, dtype, nelem):
        self.dtype = dtype if type(dtype) is _DTYPE else np.dtype(dtype)
        self.nelem = int(nelem)
        self.handle = api.new(self.dtype, max(1, self.nelem))

    def __del__(self):
        try:
            if self.handle is not None:
                api.free(self.handle)
        except Exception:
            pass


_STRIDES = {}


def _contiguous_strides(shape):
    st = _STRIDES.get(shape)
    if st is None:
        acc, s = [], 1
        for d in reversed(shape):
            acc.append(s)
            s *= d
        st = tuple(reversed(acc))
        if len(_STRIDES) < 65536:
            _STRIDES[shape] = st
    return st


class BhArray:
    """A view (offset, shape, strides in elements) into a BhBase."""

    __array_priority__ = 100

    def __init__(self, shape, dtype=np.float64, base=None, offset=0, strides=None, slides=()):
        if type(shape) is tuple:
            for s in shape:
                if type(s) is not int:
                    shape = tuple(int(s) for s in shape)
                    break
        elif isinstance(shape, numbers.Integral):
            shape = (int(shape),)
        else:
            shape = tuple(int(s) for s in shape)
        self.shape = shape
        if base is None:
            self.dtype = dtype if type(dtype) is _DTYPE else np.dtype(dtype)
            n = 1
            for s in shape:
                n *= s
            self.base = BhBase(self.dtype, n)
        else:
            self.dtype = base.dtype
            self.base = base
        self.offset = offset if type(offset) is int else int(offset)
        self.strides = _contiguous_strides(shape) if strides is None else tuple(strides)
        # Slide dims (rank, offset_change, shape_change, stride, shape, step_delay) applied between
        # the repeats of do_while (views indexed by an Iterator).
        self.slides = tuple(slides)

    @classmethod
    def _fresh(cls, shape, dtype):
        """A new contiguous array of `shape` (a tuple of ints) and `dtype` (a numpy dtype), both taken
        from an existing array, so without the constructor's conversions."""
        self = object.__new__(cls)
        self.shape = shape
        self.dtype = dtype
        n = 1
        for d in shape:
            n *= d
        self.base = BhBase(dtype, n)
        self.offset = 0
        self.strides = _contiguous_strides(shape)
        self.slides = ()
        return self

    # ------------------------------------------------------------ basics
    @property
    def ndim(self):
        return len(self.shape)

    @property
    def size(self):
        n = 1
        for s in self.shape:
            n *= s
        return n

    def __len__(self):
        return self.shape[0]

    def iscontiguous(self):
        return self.strides == _contiguous_strides(self.shape) or self.size <= 1

    def empty(self):
        return self.size == 0

    def view_desc(self):
        """(handle, start, shape, stride) of this view for the bridge; 0-d views are 1-element 1-d views."""
        if not self.shape:
            return (self.base.handle, self.offset, (1,), (1,))
        return (self.base.handle, self.offset, self.shape, self.strides)

    def __repr__(self):
        return "BhArray(%r, dtype=%s)" % (self.asnumpy(), self.dtype)

    # ------------------------------------------------------------ data
    def _small_view(self):
        """A view of at most 1/8 of a large base: fetching it copies it to a base of its own on the
        device first, so only its elements cross to the host, not the whole base."""
        return 0 < self.size and self.base.nelem >= 65536 and self.size * 8 <= self.base.nelem

    def asnumpy(self):
        """Flush and return a NumPy copy of this view."""
        if self._small_view():
            return self.copy().asnumpy()
        api.sync(self.base.handle)
        api.flush()
        host = api.data_copy_out(self.base.handle, self.dtype, max(1, self.base.nelem))
        if self.size == 0:
            return np.empty(self.shape, dtype=self.dtype)
        it = host.itemsize
        v = np.lib.stride_tricks.as_strided(host[self.offset:], shape=self.shape,
                                            strides=tuple(s * it for s in self.strides))
        return np.array(v, copy=True)

    copy2numpy = asnumpy

    def __array__(self, dtype=None, copy=None):
        a = self.asnumpy()
        return a if dtype is None else a.astype(dtype)

    def item(self):
        return self.asnumpy().item()

    def __float__(self):
        return float(self.item())

    def __int__(self):
        return int(self.item())

    def __bool__(self):
        return bool(self.item())

    # ------------------------------------------------------------ views
    def _view(self, offset, shape, strides, slides=None):
        return BhArray(shape, base=self.base, offset=offset, strides=strides,
                       slides=self.slides if slides is None else slides)

    def __getitem__(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        if any(k is Ellipsis for k in key):
            i = next(j for j, k in enumerate(key) if k is Ellipsis)
            fill = self.ndim - (len(key) - 1 - sum(1 for k in key if k is None))
            key = key[:i] + (slice(None),) * fill + key[i + 1:]
        offset, shape, strides = self.offset, [], []
        slides = list(self.slides)
        dim = 0
        for k in key:
            if k is None:
                shape.append(1)
                strides.append(0)
                continue
            if dim >= self.ndim:
                raise IndexError("too many indices")
            n, st = self.shape[dim], self.strides[dim]
            if isinstance(k, slice):
                start, stop, step = k.indices(n)
                cnt = len(range(start, stop, step))
                offset += start * st if cnt > 0 else 0
                shape.append(cnt)
                strides.append(st * step)
            elif isinstance(k, Iterator):
                offset += (k.val % n) * st
                slides.append((0, k.step, 0, st, n, k.step_delay))
            elif isinstance(k, numbers.Integral):
                k = int(k)
                if k < 0:
                    k += n
                if not 0 <= k < n:
                    raise IndexError("index out of range")
                offset += k * st
            else:
                raise TypeError("unsupported index %r" % (k,))
            dim += 1
        for d in range(dim, self.ndim):
            shape.append(self.shape[d])
            strides.append(self.strides[d])
        return self._view(offset, tuple(shape), tuple(strides), tuple(slides))

    def __setitem__(self, key, value):
        assign(value, self[key])

    def reshape(self, *shape):
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        shape = list(shape)
        if -1 in shape:
            i = shape.index(-1)
            rest = 1
            for j, s in enumerate(shape):
                if j != i:
                    rest *= s
            shape[i] = self.size // max(1, rest)
        shape = tuple(int(s) for s in shape)
        src = self if self.iscontiguous() else self.copy()
        return BhArray(shape, base=src.base, offset=src.offset, strides=_contiguous_strides(shape))

    def flatten(self, always_copy=True):
        if self.iscontiguous() and not always_copy:
            return self.reshape(self.size)
        return self.copy().reshape(self.size)

    def ravel(self):
        return self.flatten(always_copy=False)

    def transpose(self, *axes):
        if not axes:
            axes = tuple(reversed(range(self.ndim)))
        elif len(axes) == 1 and isinstance(axes[0], (tuple, list)):
            axes = tuple(axes[0])
        return self._view(self.offset, tuple(self.shape[a] for a in axes), tuple(self.strides[a] for a in axes))

    @property
    def T(self):
        return self.transpose()

    def copy(self):
        out = BhArray(self.shape, self.dtype)
        ops_identity(self, out)
        return out

    def astype(self, dtype):
        if not self.slides:
            out = BhArray._fresh(self.shape, dtype if type(dtype) is _DTYPE else np.dtype(dtype))
            _op(ops.IDENTITY, [out.view_desc(), self.view_desc()])
            return out
        out = BhArray(self.shape, dtype)
        ops_identity(self, out)
        return out

    def fill(self, value):
        assign(value, self)

    # ------------------------------------------------------------ reductions
    def sum(self, axis=None):
        return ufunc("add").reduce(self, axis)

    def prod(self, axis=None):
        return ufunc("multiply").reduce(self, axis)

    def max(self, axis=None):
        return ufunc("maximum").reduce(self, axis)

    def min(self, axis=None):
        return ufunc("minimum").reduce(self, axis)

    def any(self, axis=None):
        return ufunc("logical_or").reduce(self, axis)

    def all(self, axis=None):
        return ufunc("logical_and").reduce(self, axis)

    def mean(self, axis=None):
        s = self.sum(axis)
        n = self.size if axis is None else self.shape[axis]
        return s / n

    def cumsum(self, axis=0):
        return ufunc("add").accumulate(self, axis)

    def cumprod(self, axis=0):
        return ufunc("multiply").accumulate(self, axis)

    # ------------------------------------------------------------ operators
    def __add__(self, o): return ufunc("add")(self, o)
    def __radd__(self, o): return ufunc("add")(o, self)
    def __sub__(self, o): return ufunc("subtract")(self, o)
    def __rsub__(self, o): return ufunc("subtract")(o, self)
    def __mul__(self, o): return ufunc("multiply")(self, o)
    def __rmul__(self, o): return ufunc("multiply")(o, self)
    def __truediv__(self, o): return ufunc("true_divide")(self, o)
    def __rtruediv__(self, o): return ufunc("true_divide")(o, self)
    def __floordiv__(self, o): return ufunc("floor_divide")(self, o)
    def __rfloordiv__(self, o): return ufunc("floor_divide")(o, self)
    def __mod__(self, o): return ufunc("remainder")(self, o)
    def __rmod__(self, o): return ufunc("remainder")(o, self)
    def __pow__(self, o): return ufunc("power")(self, o)
    def __rpow__(self, o): return ufunc("power")(o, self)
    def __neg__(self): return ufunc("multiply")(self, -1)
    def __pos__(self): return self
    def __abs__(self): return ufunc("absolute")(self)
    def __invert__(self): return ufunc("invert")(self)
    def __and__(self, o): return ufunc("bitwise_and")(self, o)
    def __or__(self, o): return ufunc("bitwise_or")(self, o)
    def __xor__(self, o): return ufunc("bitwise_xor")(self, o)
    def __lshift__(self, o): return ufunc("left_shift")(self, o)
    def __rshift__(self, o): return ufunc("right_shift")(self, o)
    def __lt__(self, o): return ufunc("less")(self, o)
    def __le__(self, o): return ufunc("less_equal")(self, o)
    def __gt__(self, o): return ufunc("greater")(self, o)
    def __ge__(self, o): return ufunc("greater_equal")(self, o)
    def __eq__(self, o): return ufunc("equal")(self, o)
    def __ne__(self, o): return ufunc("not_equal")(self, o)
    __hash__ = object.__hash__

    def __iadd__(self, o): return ufunc("add")(self, o, out=self)
    def __isub__(self, o): return ufunc("subtract")(self, o, out=self)
    def __imul__(self, o): return ufunc("multiply")(self, o, out=self)
    def __itruediv__(self, o): return ufunc("true_divide")(self, o, out=self)


class Iterator:
    """An integer-like loop counter for do_while bodies: indexing an axis with it
    gives a view that moves `step` elements along that axis at every repeat
    (reference bridge/npbackend/bohrium/loop.py:287-312)."""

    def __init__(self, val=0, step=1, step_delay=1):
        self.val, self.step, self.step_delay = int(val), int(step), int(step_delay)

    def __add__(self, k):
        return Iterator(self.val + int(k), self.step, self.step_delay)

    __radd__ = __add__

    def __sub__(self, k):
        return Iterator(self.val - int(k), self.step, self.step_delay)

    def __rsub__(self, k):
        return Iterator(int(k) - self.val, -self.step, self.step_delay)

    def __mul__(self, k):
        return Iterator(self.val * int(k), self.step * int(k), self.step_delay)

    __rmul__ = __mul__

    def __neg__(self):
        return Iterator(-self.val, -self.step, self.step_delay)


def get_iterator(val=0):
    return Iterator(val, 1)


# ---------------------------------------------------------------- helpers

def _broadcast_view(a, shape):
    """View `a` broadcast to `shape` (zero strides)."""
    if a.shape == tuple(shape):
        return a
    nd = len(shape)
    pad = nd - a.ndim
    if pad < 0:
        raise ValueError("cannot broadcast %s to %s" % (a.shape, shape))
    ash = (1,) * pad + a.shape
    ast = (0,) * pad + a.strides
    strides = []
    for s_in, st, s_out in zip(ash, ast, shape):
        if s_in == s_out:
            strides.append(st)
        elif s_in == 1:
            strides.append(0)
        else:
            raise ValueError("cannot broadcast %s to %s" % (a.shape, shape))
    return BhArray(tuple(shape), base=a.base, offset=a.offset, strides=tuple(strides), slides=a.slides)


def _as_operand(x):
    return x.view_desc() if isinstance(x, BhArray) else None


def _slides_of(arrays):
    for a in arrays:
        if isinstance(a, BhArray) and a.slides:
            return [a.slides if isinstance(a, BhArray) else () for a in arrays]
    return None


def _same_memory(a, b):
    return isinstance(a, BhArray) and isinstance(b, BhArray) and a.base is b.base


def ops_identity(src, dst):
    """dst[...] = src with a C cast (BH_IDENTITY)."""
    if isinstance(src, BhArray):
        api.op(ops.IDENTITY, [dst.view_desc(), _broadcast_view(src, dst.shape).view_desc()],
               slides=_slides_of([dst, src]))
    else:
        api.op(ops.IDENTITY, [dst.view_desc(), None], const=_const_scalar(src, dst.dtype), slides=_slides_of([dst]))


_CONSTS = {}
_CONST_BYTES = {}  # constant key -> (bridge type code, bytes), for ufunc._call_direct
# ufunc calls take ufunc._call_direct when it applies, other calls _op's direct form (a switch for
# tests that compare both paths)
_DIRECT = [True]
_AXES = {}  # axis -> its constant (int64 scalar)


def _op(opcode, descs, const=None, slides=None):
    """api.op, straight into the bridge extension when nothing needs the general path (no slides, no
    trace or recording)."""
    if _DIRECT[0] and not slides and api.direct_ready():
        if const is None:
            api.op_direct(opcode, descs, -1, None)
        else:
            api.op_direct(opcode, descs, api.BH_TYPES[const.dtype], const.tobytes())
        return
    api.op(opcode, descs, const=const, slides=slides)


def _axis_const(axis):
    c = _AXES.get(axis)
    if c is None:
        c = _AXES[axis] = np.int64(axis)
    return c


def _const_key(value, dtype):
    # 0.0 == -0.0 as dictionary keys: zeros are told apart by their text
    return (type(value), value if value or type(value) is not float else repr(value), dtype)


def _const_scalar(value, dtype):
    """A constant in the instruction's type, converted the way the reference's bridges do (C cast)."""
    if type(value) in (int, float, bool):
        key = _const_key(value, dtype)
        c = _CONSTS.get(key)
        if c is None:
            c = np.asarray(value).astype(np.dtype(dtype))[()]
            if len(_CONSTS) < 65536:
                _CONSTS[key] = c
        return c
    return np.asarray(value).astype(np.dtype(dtype))[()]


def assign(value, out):
    if isinstance(value, BhArray):
        if value.dtype == out.dtype and value.shape == out.shape and value.base is out.base and \
                value.offset == out.offset and value.strides == out.strides:
            return out
        if _same_memory(value, out):
            value = value.copy()
        ops_identity(value, out)
    else:
        ops_identity(value, out)
    return out


def array(obj, dtype=None):
    """Create a BhArray holding a copy of `obj` (anything numpy.asarray accepts)."""
    a = np.asarray(obj, dtype=dtype)
    if a.dtype not in api.BH_TYPES:
        a = a.astype(np.float64)
    out = BhArray(a.shape, a.dtype)
    if out.size:
        api.data_copy_in(out.base.handle, np.ascontiguousarray(a).reshape(-1))
    return out


def empty(shape, dtype=np.float64):
    return BhArray(shape, dtype)


def full(shape, value, dtype=np.float64):
    out = BhArray(shape, dtype)
    if out.size:
        ops_identity(value, out)
    return out


def zeros(shape, dtype=np.float64):
    return full(shape, 0, dtype)


def ones(shape, dtype=np.float64):
    return full(shape, 1, dtype)


def empty_like(a, dtype=None):
    return empty(a.shape, a.dtype if dtype is None else dtype)


def zeros_like(a, dtype=None):
    return zeros(a.shape, a.dtype if dtype is None else dtype)


def ones_like(a, dtype=None):
    return ones(a.shape, a.dtype if dtype is None else dtype)


def arange(start, stop=None, step=1, dtype=None):
    """Evenly spaced values (BH_RANGE, then an affine map when start/step are not 0/1)."""
    if stop is None:
        start, stop = 0, start
    n = max(0, int(np.ceil((stop - start) / step)))
    if dtype is None:
        dtype = np.result_type(np.asarray(start), np.asarray(stop), np.asarray(step))
        if dtype.kind in "iub":
            dtype = np.dtype(np.int64)
    dtype = np.dtype(dtype)
    out = BhArray((n,), np.int64 if dtype.kind in "iu" else dtype)
    if n > 0:
        api.op(ops.RANGE, [out.view_desc()])
        if step != 1:
            out = out * step
        if start != 0:
            out = out + start
    return out if out.dtype == dtype else out.astype(dtype)


def flush():
    api.flush()


def asnumpy_many(arrays):
    """Fetch several arrays with a single flush (small views of large bases through copies of their own)."""
    arrays = [a.copy() if a._small_view() else a for a in arrays]
    for a in arrays:
        api.sync(a.base.handle)
    api.flush()
    return [a.asnumpy() for a in arrays]


def take(a, indices):
    """out[i] = a.flat[indices[i]] for a contiguous `a` (BH_GATHER)."""
    if not a.iscontiguous():
        a = a.copy()
    idx = indices if isinstance(indices, BhArray) else array(indices, np.int64)
    if idx.dtype != np.int64:
        idx = idx.astype(np.int64)
    out = BhArray._fresh(idx.shape, a.dtype)
    _op(ops.GATHER, [out.view_desc(), (a.base.handle, a.offset, (a.size,), (1,)), idx.view_desc()])
    return out


def put(a, indices, values, mask=None):
    """a.flat[indices[i]] = values[i] (BH_SCATTER), only where mask[i] if given (BH_COND_SCATTER)."""
    idx = indices if isinstance(indices, BhArray) else array(indices, np.int64)
    if idx.dtype != np.int64:
        idx = idx.astype(np.int64)
    vals = values if isinstance(values, BhArray) else full(idx.shape, values, a.dtype)
    if vals.dtype != a.dtype:
        vals = vals.astype(a.dtype)
    vals = _broadcast_view(vals, idx.shape)
    flat = BhArray((a.size,), base=a.base, offset=a.offset)
    if not a.iscontiguous():
        raise ValueError("put() needs a contiguous target")
    if mask is None:
        api.op(ops.SCATTER, [flat.view_desc(), vals.view_desc(), idx.view_desc()])
    else:
        m = mask if isinstance(mask, BhArray) else array(mask, np.bool_)
        api.op(ops.COND_SCATTER, [flat.view_desc(), vals.view_desc(), idx.view_desc(),
                                  _broadcast_view(m, idx.shape).view_desc()])


# ---------------------------------------------------------------- ufuncs

_UFUNCS = {}


def ufunc(name):
    return _UFUNCS[name]


class Ufunc:
    """An elementwise operation mapped onto one opcode; dtypes are resolved as NumPy does."""

    def __init__(self, name, opcode, nin, np_name=None, reduce_opcode=None, accumulate_opcode=None):
        self.name = name
        self.opcode = opcode
        self.nin = nin
        self.np_ufunc = getattr(np, np_name or name)
        self.reduce_opcode = reduce_opcode
        self.accumulate_opcode = accumulate_opcode
        self._cache = {}
        _UFUNCS[name] = self

    def _loop_dtypes(self, args):
        key = tuple(a.dtype if isinstance(a, BhArray) else type(a) if isinstance(a, (bool, int, float, complex))
                    else getattr(a, "dtype", None) for a in args)
        hit = self._cache.get(key)
        if hit is not None:
            return hit
        res = self._resolve(args)
        self._cache[key] = res
        return res

    def _resolve(self, args):
        types = []
        for a in args:
            if isinstance(a, BhArray):
                types.append(a.dtype)
            elif isinstance(a, (bool, np.bool_)):
                types.append(np.dtype(np.bool_))
            elif isinstance(a, np.generic):
                types.append(a.dtype)
            elif isinstance(a, numbers.Integral):
                types.append(int)
            elif isinstance(a, numbers.Real):
                types.append(float)
            elif isinstance(a, numbers.Complex):
                types.append(complex)
            else:
                types.append(np.asarray(a).dtype)
        if self.name in ("real", "imag"):
            t = np.dtype(types[0]) if not isinstance(types[0], type) else np.dtype(np.float64)
            out = np.dtype(np.float32) if t == np.complex64 else np.dtype(np.float64) if t.kind == "c" else t
            return (t,), out
        res = self.np_ufunc.resolve_dtypes(tuple(types) + (None,))
        return res[:self.nin], res[self.nin]

    def __call__(self, *args, out=None):
        if out is None and _DIRECT[0] and api.direct_ready():
            res = self._call_direct(args)
            if res is not None:
                return res
        if len(args) != self.nin:
            raise TypeError("%s takes %d inputs" % (self.name, self.nin))
        in_types, out_type = self._loop_dtypes(args)
        shapes = [a.shape for a in args if isinstance(a, BhArray)]
        if not shapes:
            shape = ()
        elif len(shapes) == 1 or all(sh == shapes[0] for sh in shapes):
            shape = shapes[0]
        else:
            shape = np.broadcast_shapes(*shapes)
        if out is not None:
            shape = out.shape
        operands = []
        const = None
        for a, t in zip(args, in_types):
            if isinstance(a, BhArray):
                if a.dtype != t:
                    a = a.astype(t)
                operands.append(a if a.shape == shape else _broadcast_view(a, shape))
            else:
                if const is not None:
                    a = full(shape, a, t)
                    operands.append(a)
                    continue
                const = _const_scalar(a, t)
                operands.append(None)
        if const is not None and all(o is None for o in operands):
            operands[0] = full(shape, args[0], in_types[0])
            const = None if self.nin == 1 else const
        res = out if (out is not None and out.dtype == out_type) else BhArray(shape, out_type)
        self._emit(res, operands, const)
        if out is not None and res is not out:
            ops_identity(res, out)
            return out
        return res

    def _call_direct(self, args):
        """The common calls without the general machinery: array operands of one shape, or an array and
        a Python int or float, no slides, loop types equal to the arrays' own (no cast); the instruction
        goes straight into the extension.  None for every other call."""
        op = self.opcode
        if op is None or callable(op) or len(args) != self.nin:
            return None
        a = args[0]
        if type(a) is not BhArray or a.slides:
            return None
        if self.nin == 1:
            hit = self._cache.get((a.dtype,)) or self._loop_dtypes(args)
            if hit[0][0] != a.dtype:
                return None
            res = BhArray._fresh(a.shape, hit[1])
            api.op_direct(op, [res.view_desc(), a.view_desc()], -1, None)
            return res
        b = args[1]
        tb = type(b)
        if tb is BhArray:
            if b.slides or b.shape != a.shape:
                return None
            hit = self._cache.get((a.dtype, b.dtype)) or self._loop_dtypes(args)
            if hit[0][0] != a.dtype or hit[0][1] != b.dtype:
                return None
            res = BhArray._fresh(a.shape, hit[1])
            api.op_direct(op, [res.view_desc(), a.view_desc(), b.view_desc()], -1, None)
            return res
        if tb is not float and tb is not int:
            return None
        hit = self._cache.get((a.dtype, tb)) or self._loop_dtypes(args)
        if hit[0][0] != a.dtype:
            return None
        key = _const_key(b, hit[0][1])
        cb = _CONST_BYTES.get(key)
        if cb is None:
            c = _const_scalar(b, hit[0][1])
            cb = (api.BH_TYPES[c.dtype], c.tobytes())
            if len(_CONST_BYTES) < 65536:
                _CONST_BYTES[key] = cb
        res = BhArray._fresh(a.shape, hit[1])
        api.op_direct(op, [res.view_desc(), a.view_desc(), None], cb[0], cb[1])
        return res

    def _emit(self, res, operands, const):
        descs = [res.view_desc()] + [_as_operand(o) for o in operands]
        if self.opcode is None:
            raise NotImplementedError(self.name)
        if callable(self.opcode):
            self.opcode(res, operands, const)
            return
        api.op(self.opcode, descs, const=const, slides=_slides_of([res] + list(operands)))

    # -------------------------------------------------------- reduce
    def reduce(self, ary, axis=0, out=None):
        if self.reduce_opcode is None:
            raise NotImplementedError("%s.reduce" % self.name)
        if not isinstance(ary, BhArray):
            ary = array(ary)
        if axis is None:
            axis = list(range(ary.ndim))
        elif isinstance(axis, numbers.Integral):
            axis = [axis]
        axis = sorted(a + ary.ndim if a < 0 else a for a in axis)
        if not self.name.startswith("logical") and ary.dtype == np.bool_:
            ary = ary.astype(np.uint64)  # counting True values, as the reference's bh107
        if ary.ndim == 0:
            return ary.copy()
        if len(axis) > 1:
            if len(axis) == ary.ndim and ary.iscontiguous():
                return self.reduce(ary.reshape(ary.size), 0, out)
            r = self.reduce(ary, axis[-1])
            return self.reduce(r, axis[:-1], out)
        ax = axis[0]
        shape = tuple(s for i, s in enumerate(ary.shape) if i != ax)
        tmp = BhArray._fresh(shape, ary.dtype)
        if ary.shape[ax] == 0:
            ident = {"add": 0, "multiply": 1, "logical_or": False, "logical_and": True}.get(self.name)
            if ident is None:
                raise ValueError("zero-size reduction without identity")
            ops_identity(ident, tmp)
        elif ary.shape[ax] == 1:
            ops_identity(ary[(slice(None),) * ax + (0,)], tmp)
        else:
            _op(self.reduce_opcode, [tmp.view_desc(), ary.view_desc(), None], const=_axis_const(ax),
                slides=_slides_of([tmp, ary]))
        if out is not None:
            assign(tmp, out)
            return out
        return tmp

    def accumulate(self, ary, axis=0, out=None):
        if self.accumulate_opcode is None:
            raise NotImplementedError("%s.accumulate" % self.name)
        if not isinstance(ary, BhArray):
            ary = array(ary)
        if axis < 0:
            axis += ary.ndim
        dt = np.dtype(np.int64) if ary.dtype.kind in "biu" and ary.dtype.itemsize < 8 else ary.dtype
        if ary.dtype == np.uint64:
            dt = np.dtype(np.uint64)
        if dt != ary.dtype:
            ary = ary.astype(dt)
        res = BhArray._fresh(ary.shape, dt)
        if ary.size:
            _op(self.accumulate_opcode, [res.view_desc(), ary.view_desc(), None], const=_axis_const(axis))
        if out is not None:
            assign(res, out)
            return out
        return res


def _floor_divide(res, operands, const):
    """Integer: BH_DIVIDE (floor semantics for signed types); float: floor(a / b)."""
    descs = [res.view_desc()] + [_as_operand(o) for o in operands]
    api.op(ops.DIVIDE, descs, const=const)
    if res.dtype.kind in "fc":
        api.op(ops.FLOOR, [res.view_desc(), res.view_desc()])


Ufunc("add", ops.ADD, 2, reduce_opcode=ops.ADD_REDUCE, accumulate_opcode=ops.ADD_ACCUMULATE)
Ufunc("subtract", ops.SUBTRACT, 2)
Ufunc("multiply", ops.MULTIPLY, 2, reduce_opcode=ops.MULTIPLY_REDUCE, accumulate_opcode=ops.MULTIPLY_ACCUMULATE)
Ufunc("true_divide", ops.DIVIDE, 2, np_name="true_divide")
Ufunc("divide", ops.DIVIDE, 2, np_name="true_divide")
Ufunc("floor_divide", _floor_divide, 2)
Ufunc("power", ops.POWER, 2)
Ufunc("absolute", ops.ABSOLUTE, 1)
Ufunc("greater", ops.GREATER, 2)
Ufunc("greater_equal", ops.GREATER_EQUAL, 2)
Ufunc("less", ops.LESS, 2)
Ufunc("less_equal", ops.LESS_EQUAL, 2)
Ufunc("equal", ops.EQUAL, 2)
Ufunc("not_equal", ops.NOT_EQUAL, 2)
Ufunc("logical_and", ops.LOGICAL_AND, 2, reduce_opcode=ops.LOGICAL_AND_REDUCE)
Ufunc("logical_or", ops.LOGICAL_OR, 2, reduce_opcode=ops.LOGICAL_OR_REDUCE)
Ufunc("logical_xor", ops.LOGICAL_XOR, 2, reduce_opcode=ops.LOGICAL_XOR_REDUCE)
Ufunc("logical_not", ops.LOGICAL_NOT, 1)
Ufunc("maximum", ops.MAXIMUM, 2, reduce_opcode=ops.MAXIMUM_REDUCE)
Ufunc("minimum", ops.MINIMUM, 2, reduce_opcode=ops.MINIMUM_REDUCE)
Ufunc("bitwise_and", ops.BITWISE_AND, 2, reduce_opcode=ops.BITWISE_AND_REDUCE)
Ufunc("bitwise_or", ops.BITWISE_OR, 2, reduce_opcode=ops.BITWISE_OR_REDUCE)
Ufunc("bitwise_xor", ops.BITWISE_XOR, 2, reduce_opcode=ops.BITWISE_XOR_REDUCE)
Ufunc("invert", ops.INVERT, 1)
Ufunc("left_shift", ops.LEFT_SHIFT, 2)
Ufunc("right_shift", ops.RIGHT_SHIFT, 2)
for _n, _o in [("cos", ops.COS), ("sin", ops.SIN), ("tan", ops.TAN), ("cosh", ops.COSH), ("sinh", ops.SINH),
               ("tanh", ops.TANH), ("arcsin", ops.ARCSIN), ("arccos", ops.ARCCOS), ("arctan", ops.ARCTAN),
               ("arcsinh", ops.ARCSINH), ("arccosh", ops.ARCCOSH), ("arctanh", ops.ARCTANH), ("exp", ops.EXP),
               ("exp2", ops.EXP2), ("expm1", ops.EXPM1), ("log", ops.LOG), ("log2", ops.LOG2),
               ("log10", ops.LOG10), ("log1p", ops.LOG1P), ("sqrt", ops.SQRT), ("ceil", ops.CEIL),
               ("trunc", ops.TRUNC), ("floor", ops.FLOOR), ("rint", ops.RINT), ("isnan", ops.ISNAN),
               ("isinf", ops.ISINF), ("isfinite", ops.ISFINITE), ("sign", ops.SIGN), ("real", ops.REAL),
               ("imag", ops.IMAG), ("conjugate", ops.CONJ)]:
    Ufunc(_n, _o, 1)
Ufunc("arctan2", ops.ARCTAN2, 2)
Ufunc("fmod", ops.MOD, 2)
Ufunc("remainder", ops.REMAINDER, 2)
Ufunc("mod", ops.REMAINDER, 2, np_name="remainder")
