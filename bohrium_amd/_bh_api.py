"""ctypes binding of the bhc C-ABI (include/bhc.h).

The function names follow the reference's low-level Python API `_bh_api`
(bridge/py_api/src/_bh_api.c:617-630: new, op, flush, data_get, random123 ...).
Every call can optionally be recorded into a trace: a list of plain tuples
describing exactly what was sent to the bridge, which a checker can replay.
"""
import atexit
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")
CONFIG = os.path.join(_HERE, "config.ini")
MAXDIM = 16

# bh_type ids (reference core/codegen/types.json)
BH_TYPES = {
    np.dtype(np.bool_): 0, np.dtype(np.int8): 1, np.dtype(np.int16): 2, np.dtype(np.int32): 3,
    np.dtype(np.int64): 4, np.dtype(np.uint8): 5, np.dtype(np.uint16): 6, np.dtype(np.uint32): 7,
    np.dtype(np.uint64): 8, np.dtype(np.float32): 9, np.dtype(np.float64): 10,
    np.dtype(np.complex64): 11, np.dtype(np.complex128): 12,
}
R123 = 13


class BhcView(ctypes.Structure):
    _fields_ = [("base", ctypes.c_void_p), ("start", ctypes.c_int64), ("ndim", ctypes.c_int64),
                ("shape", ctypes.c_int64 * MAXDIM), ("stride", ctypes.c_int64 * MAXDIM)]


class BohriumError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()
# Record-only mode: calls are traced but nothing reaches the native bridge
# (lets a checker build traces on a machine without a GPU).
_record_only = [False]
_fake_handle = [0]
_fake_handles = set()
_trace = None          # list of records while tracing, else None
_base_ids = {}         # handle -> small integer id used in traces
_next_base_id = [0]


def _load():
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(LIBDIR, "libbhc.so")
    if not os.path.exists(path):
        raise BohriumError("bohrium_amd native libraries are not built (%s missing); run bohrium_amd/build.py" % path)
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    vp, i64, u64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int
    lib.bhc_init.argtypes = [ctypes.c_char_p]
    lib.bhc_last_error.restype = ctypes.c_char_p
    lib.bhc_new.argtypes = [i32, i64]
    lib.bhc_new.restype = vp
    lib.bhc_op.argtypes = [i64, i32, ctypes.POINTER(BhcView), i32, vp]
    lib.bhc_random123.argtypes = [ctypes.POINTER(BhcView), u64, u64]
    lib.bhc_free.argtypes = [vp]
    lib.bhc_sync.argtypes = [vp]
    lib.bhc_flush_and_repeat.argtypes = [u64, vp]
    lib.bhc_data_get.argtypes = [vp, i32, i32, i32]
    lib.bhc_data_get.restype = vp
    lib.bhc_data_set.argtypes = [vp, i32, vp]
    lib.bhc_data_copy_in.argtypes = [vp, vp, i64]
    lib.bhc_data_copy_out.argtypes = [vp, vp, i64]
    lib.bhc_message.argtypes = [ctypes.c_char_p]
    lib.bhc_message.restype = ctypes.c_char_p
    lib.bhc_extmethod.argtypes = [ctypes.c_char_p, i32, ctypes.POINTER(BhcView)]
    lib.bhc_user_kernel.argtypes = [ctypes.c_char_p, i32, ctypes.POINTER(BhcView), ctypes.c_char_p,
                                    ctypes.c_char_p, ctypes.c_char_p]
    lib.bhc_user_kernel.restype = ctypes.c_char_p
    lib.bhc_set_flush_threshold.argtypes = [i64]
    lib.bhc_op_slide.argtypes = [i32, i64, i64, i64, i64, i64, i64]
    lib.bhc_base_nelem.argtypes = [vp]
    lib.bhc_base_nelem.restype = i64
    lib.bhc_record_flushes.argtypes = [ctypes.c_char_p]
    lib.bhc_replay_file.argtypes = [ctypes.c_char_p]
    lib.bhc_replay_base.argtypes = [u64]
    lib.bhc_replay_base.restype = vp
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise BohriumError(_lib.bhc_last_error().decode(errors="replace"))


_initialised = [False]
try:
    from . import _bhc as _fast  # CPython extension over libbhc (built by build.py)
except ImportError:  # pragma: no cover - the ctypes path is complete on its own
    _fast = None


def init():
    """Load the component stack (BH_STACK; the packaged config.ini is the fallback)."""
    with _lock:
        if _initialised[0]:
            return
        lib = _load()
        _check(lib.bhc_init(CONFIG.encode()))
        _initialised[0] = True
        atexit.register(shutdown)


def shutdown():
    """Flush what is queued and destroy the component stack (device work drains first)."""
    if _initialised[0]:
        _initialised[0] = False
        _lib.bhc_shutdown()


def lib():
    if not _initialised[0]:
        init()
    return _lib


class record_only:
    """Context manager: trace bridge calls without executing them."""

    def __enter__(self):
        self._prev = _record_only[0]
        _record_only[0] = True
        trace_start()
        return self

    def __exit__(self, *exc):
        self.trace = trace_stop()
        _record_only[0] = self._prev
        return False


# ---------------------------------------------------------------- tracing

def trace_start():
    """Start recording every bridge call (see oracle.replay for the format)."""
    global _trace
    _trace = []


def trace_stop():
    global _trace
    t, _trace = _trace, None
    return t


def _bid(handle):
    """The trace id of a live base.  Ids are never reused; a handle (the runtime's base address)
    can be, once its base is freed, and then names a new base with a new id (see _retire)."""
    if handle not in _base_ids:
        _base_ids[handle] = _next_base_id[0]
        _next_base_id[0] += 1
    return _base_ids[handle]


def _retire(handle):
    _base_ids.pop(handle, None)


def _rec(*record):
    if _trace is not None:
        _trace.append(record)


# ---------------------------------------------------------------- API

def new(dtype, nelem):
    if _fast is not None and _trace is None and _initialised[0] and not _record_only[0] and type(nelem) is int:
        # the common case: one call into the extension
        try:
            return _fast.new(BH_TYPES[dtype if isinstance(dtype, np.dtype) else np.dtype(dtype)], nelem)
        except RuntimeError as e:
            raise BohriumError(str(e))
    if not isinstance(dtype, np.dtype):
        dtype = np.dtype(dtype)
    if _record_only[0]:
        _fake_handle[0] += 1
        h = ("fake", _fake_handle[0])
        _fake_handles.add(h)
        _rec("new", _bid(h), dtype.str, int(nelem))
        return h
    lib()
    if _fast is not None:
        try:
            h = _fast.new(BH_TYPES[dtype], int(nelem))
        except RuntimeError as e:
            raise BohriumError(str(e))
    else:
        h = _lib.bhc_new(BH_TYPES[dtype], int(nelem))
        if not h:
            raise BohriumError(_lib.bhc_last_error().decode())
    if _trace is not None:
        _rec("new", _bid(h), dtype.str, int(nelem))
    return h


def direct_ready():
    """True when instructions can go straight into the extension: no trace, no recording."""
    return _fast is not None and _trace is None and _initialised[0] and not _record_only[0]


def op_direct(opcode, operands, ctype, cbytes):
    """Queue one instruction straight into the extension (callers check direct_ready()): `operands`
    as for op(), the constant as its bridge type code (-1: none) and its bytes."""
    try:
        _fast.op(opcode, operands, ctype, cbytes)
    except RuntimeError as e:
        raise BohriumError(str(e))


def make_view(handle, start, shape, stride):
    v = BhcView()
    v.base = handle
    v.start = int(start)
    if len(shape) > MAXDIM:
        raise BohriumError("at most %d dimensions" % MAXDIM)
    v.ndim = len(shape)
    for i, (s, st) in enumerate(zip(shape, stride)):
        v.shape[i] = int(s)
        v.stride[i] = int(st)
    return v


def op(opcode, operands, const=None, slides=None):
    """Queue one instruction.  `operands` are (handle, start, shape, stride)
    tuples or None for the constant operand, whose value is `const`
    (a numpy scalar of the constant's type).  `slides` optionally gives, per
    operand, a tuple of slide dims (rank, offset_change, shape_change, stride,
    shape, step_delay) that move the view between repeats."""
    if _fast is not None and _trace is None and not slides and _initialised[0] and not _record_only[0]:
        # the common case: one call into the extension
        try:
            if const is None:
                _fast.op(opcode, operands, -1, None)
            else:
                c = const if isinstance(const, np.generic) else np.asarray(const)
                _fast.op(opcode, operands, BH_TYPES[c.dtype], c.tobytes())
        except RuntimeError as e:
            raise BohriumError(str(e))
        return
    ctype, cbytes = -1, None
    if const is not None:
        c = const if isinstance(const, np.generic) else np.asarray(const)
        ctype = BH_TYPES[c.dtype]
        cbytes = c.tobytes()
    if not _record_only[0]:
        lib()
        if _fast is not None:
            try:
                _fast.op(int(opcode), operands, ctype, cbytes)
            except RuntimeError as e:
                raise BohriumError(str(e))
        else:
            n = len(operands)
            arr = (BhcView * max(1, n))()
            for i, o in enumerate(operands):
                arr[i] = BhcView() if o is None else make_view(*o)
            keep = None if cbytes is None else ctypes.create_string_buffer(cbytes, 16)
            _check(_lib.bhc_op(int(opcode), n, arr, ctype, keep))
        if slides:
            for k, dims in enumerate(slides):
                for d in (dims or ()):
                    if _fast is not None:
                        _fast.op_slide(k, *[int(x) for x in d])
                    else:
                        _check(_lib.bhc_op_slide(k, *[ctypes.c_int64(int(x)) for x in d]))
    if _trace is not None:
        rec_ops = [None if o is None else (_bid(o[0]), int(o[1]), tuple(int(x) for x in o[2]),
                                           tuple(int(x) for x in o[3])) for o in operands]
        _rec("op", int(opcode), rec_ops, None if const is None else (np.asarray(const).dtype.str,
                                                                     np.asarray(const).item()),
             None if not slides else tuple(tuple(tuple(int(x) for x in d) for d in (dims or ())) for dims in slides))


def random123(handle, start, shape, stride, r_start, r_key):
    if not _record_only[0]:
        lib()
        if _fast is not None:
            try:
                _fast.random123((handle, start, tuple(shape), tuple(stride)), int(r_start), int(r_key))
            except RuntimeError as e:
                raise BohriumError(str(e))
        else:
            v = make_view(handle, start, shape, stride)
            _check(_lib.bhc_random123(ctypes.byref(v), ctypes.c_uint64(int(r_start)), ctypes.c_uint64(int(r_key))))
    _rec("op", 71, [(_bid(handle), int(start), tuple(shape), tuple(stride)), None], ("r123", (int(r_start), int(r_key))))


def free(handle):
    if _fast is not None and _trace is None and _initialised[0] and not _fake_handles:
        try:
            _fast.free(handle)
        except RuntimeError as e:
            raise BohriumError(str(e))
        if _base_ids:
            _retire(handle)
        return
    if handle in _fake_handles:
        _fake_handles.discard(handle)
        if _trace is not None:
            _rec("free", _bid(handle))
        _retire(handle)
        return
    if not _initialised[0]:
        return
    if _fast is not None:
        try:
            _fast.free(handle)
        except RuntimeError as e:
            raise BohriumError(str(e))
    else:
        _check(_lib.bhc_free(handle))
    if _trace is not None:
        _rec("free", _bid(handle))
    _retire(handle)


def sync(handle):
    if _record_only[0]:
        return
    _check(lib().bhc_sync(handle))


def flush():
    if not _record_only[0]:
        _check(lib().bhc_flush())
    _rec("flush", 1, None)


def flush_and_repeat(nrepeats, cond_handle=None):
    if not _record_only[0]:
        _check(lib().bhc_flush_and_repeat(int(nrepeats), cond_handle))
    _rec("flush", int(nrepeats), None if cond_handle is None else _bid(cond_handle))


def data_copy_in(handle, array):
    a = np.ascontiguousarray(array)
    if not _record_only[0]:
        _check(lib().bhc_data_copy_in(handle, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
    _rec("data", _bid(handle), a.copy())


def data_copy_out(handle, dtype, nelem):
    if _record_only[0]:
        raise BohriumError("record-only mode: no data on the host")
    out = np.empty(int(nelem), dtype=dtype)
    _check(lib().bhc_data_copy_out(handle, out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
    return out


def data_get(handle, copy2host=True, force_alloc=True, nullify=False):
    p = lib().bhc_data_get(handle, int(copy2host), int(force_alloc), int(nullify))
    if not p and _lib.bhc_last_error():
        raise BohriumError(_lib.bhc_last_error().decode())
    return p


def message(msg):
    r = lib().bhc_message(msg.encode())
    if r is None:
        raise BohriumError(_lib.bhc_last_error().decode())
    return r.decode()


def extmethod(name, operands):
    n = len(operands)
    arr = (BhcView * n)()
    for i, (h, start, shape, stride) in enumerate(operands):
        arr[i] = make_view(h, start, shape, stride)
    _check(lib().bhc_extmethod(name.encode(), n, arr))
    _rec("ext", name, [(_bid(h), int(s), tuple(sh), tuple(st)) for (h, s, sh, st) in operands])


def user_kernel(kernel, operands, compile_cmd, tag, param):
    """Run a user kernel on `operands` ((handle, start, shape, stride) views); returns the component's
    message ("" on success)."""
    n = len(operands)
    arr = (BhcView * max(n, 1))()
    for i, (h, start, shape, stride) in enumerate(operands):
        arr[i] = make_view(h, start, shape, stride)
    r = lib().bhc_user_kernel(kernel.encode(), n, arr, compile_cmd.encode(), tag.encode(), param.encode())
    if r is None:
        raise BohriumError(_lib.bhc_last_error().decode(errors="replace"))
    return r.decode(errors="replace")


def set_flush_threshold(n):
    if _record_only[0]:
        return  # nothing reaches the native runtime
    _check(lib().bhc_set_flush_threshold(int(n)))


def record_flushes(path):
    """Append every later flush to the file `path` in the BhIR wire format (None stops recording).
    A base of a recorded flush is named in the file by its handle value (base_address)."""
    _check(lib().bhc_record_flushes(None if path is None else os.fsencode(path)))


def base_address(handle):
    """The address that names `handle`'s base in recorded flushes."""
    return int(handle)


def replay_file(path):
    """Execute the flushes recorded in `path` (see record_flushes) through the component stack."""
    _check(lib().bhc_replay_file(os.fsencode(path)))


def replay_base(address):
    """Handle of the base that stands in for the recorded base at `address` after replay_file
    (owned by the replay: read it, never free it)."""
    h = lib().bhc_replay_base(ctypes.c_uint64(int(address)))
    if not h:
        raise BohriumError(_lib.bhc_last_error().decode(errors="replace"))
    return h
