/* _bhc: CPython extension with the bridge calls a Python front end makes once
 * per array operation (new, op, random123, free).  It plays the role of the
 * reference's C module bohrium_api._bh_api (bridge/py_api/src/_bh_api.c:47-630)
 * and calls the bhc C-ABI directly, so queueing an instruction costs about a
 * microsecond instead of a ctypes round trip. */
#define PY_SSIZE_T_CLEAN
#include <Python.h>

#include <cstring>

#include "../../../include/bhc.h"

namespace {

PyObject *error(const char *fallback) {
    const char *msg = bhc_last_error();
    PyErr_SetString(PyExc_RuntimeError, (msg != nullptr && *msg) ? msg : fallback);
    return nullptr;
}

bool parse_view(PyObject *o, bhc_view_t *v) {
    if (o == Py_None) {
        std::memset(v, 0, sizeof(*v));
        return true;
    }
    PyObject *h, *start, *shape, *stride;
    if (!PyTuple_Check(o) || PyTuple_GET_SIZE(o) != 4) {
        PyErr_SetString(PyExc_TypeError, "operand must be None or (handle, start, shape, stride)");
        return false;
    }
    h = PyTuple_GET_ITEM(o, 0);
    start = PyTuple_GET_ITEM(o, 1);
    shape = PyTuple_GET_ITEM(o, 2);
    stride = PyTuple_GET_ITEM(o, 3);
    v->base = PyLong_AsVoidPtr(h);
    v->start = PyLong_AsLongLong(start);
    PyObject *sh = PySequence_Fast(shape, "shape must be a sequence");
    PyObject *st = PySequence_Fast(stride, "stride must be a sequence");
    if (sh == nullptr || st == nullptr) {
        Py_XDECREF(sh);
        Py_XDECREF(st);
        return false;
    }
    const Py_ssize_t nd = PySequence_Fast_GET_SIZE(sh);
    if (nd > BHC_MAXDIM || nd != PySequence_Fast_GET_SIZE(st)) {
        Py_DECREF(sh);
        Py_DECREF(st);
        PyErr_SetString(PyExc_ValueError, "bad view rank");
        return false;
    }
    v->ndim = nd;
    for (Py_ssize_t i = 0; i < nd; ++i) {
        v->shape[i] = PyLong_AsLongLong(PySequence_Fast_GET_ITEM(sh, i));
        v->stride[i] = PyLong_AsLongLong(PySequence_Fast_GET_ITEM(st, i));
    }
    Py_DECREF(sh);
    Py_DECREF(st);
    return !PyErr_Occurred();
}

PyObject *py_new(PyObject *, PyObject *args) {
    int dtype;
    long long nelem;
    if (!PyArg_ParseTuple(args, "iL", &dtype, &nelem)) return nullptr;
    bhc_base_t b = bhc_new(dtype, nelem);
    if (b == nullptr) return error("bhc_new failed");
    return PyLong_FromVoidPtr(b);
}

// op(opcode, operands_tuple, const_type, const_bytes_or_None)
PyObject *py_op(PyObject *, PyObject *args) {
    long long opcode;
    PyObject *ops;
    int ctype;
    PyObject *cbytes;
    if (!PyArg_ParseTuple(args, "LOiO", &opcode, &ops, &ctype, &cbytes)) return nullptr;
    PyObject *seq = PySequence_Fast(ops, "operands must be a sequence");
    if (seq == nullptr) return nullptr;
    const Py_ssize_t n = PySequence_Fast_GET_SIZE(seq);
    bhc_view_t views[8];
    if (n > 8) {
        Py_DECREF(seq);
        PyErr_SetString(PyExc_ValueError, "too many operands");
        return nullptr;
    }
    for (Py_ssize_t i = 0; i < n; ++i) {
        if (!parse_view(PySequence_Fast_GET_ITEM(seq, i), &views[i])) {
            Py_DECREF(seq);
            return nullptr;
        }
    }
    Py_DECREF(seq);
    const void *cval = nullptr;
    char buf[16] = {0};
    if (cbytes != Py_None) {
        char *p;
        Py_ssize_t len;
        if (PyBytes_AsStringAndSize(cbytes, &p, &len) != 0) return nullptr;
        std::memcpy(buf, p, len > 16 ? 16 : len);
        cval = buf;
    }
    if (bhc_op(opcode, static_cast<int>(n), views, ctype, cval) != 0) return error("bhc_op failed");
    Py_RETURN_NONE;
}

PyObject *py_random123(PyObject *, PyObject *args) {
    PyObject *view;
    unsigned long long start, key;
    if (!PyArg_ParseTuple(args, "OKK", &view, &start, &key)) return nullptr;
    bhc_view_t v;
    if (!parse_view(view, &v)) return nullptr;
    if (bhc_random123(&v, start, key) != 0) return error("bhc_random123 failed");
    Py_RETURN_NONE;
}

PyObject *py_op_slide(PyObject *, PyObject *args) {
    int operand;
    long long rank, off, shape_change, stride, shape, delay;
    if (!PyArg_ParseTuple(args, "iLLLLLL", &operand, &rank, &off, &shape_change, &stride, &shape, &delay))
        return nullptr;
    if (bhc_op_slide(operand, rank, off, shape_change, stride, shape, delay) != 0) return error("bhc_op_slide failed");
    Py_RETURN_NONE;
}

PyObject *py_free(PyObject *, PyObject *args) {
    PyObject *h;
    if (!PyArg_ParseTuple(args, "O", &h)) return nullptr;
    if (bhc_free(PyLong_AsVoidPtr(h)) != 0) return error("bhc_free failed");
    Py_RETURN_NONE;
}

PyMethodDef kMethods[] = {
    {"new", py_new, METH_VARARGS, "new(dtype_id, nelem) -> handle"},
    {"op", py_op, METH_VARARGS, "op(opcode, operands, const_type, const_bytes)"},
    {"random123", py_random123, METH_VARARGS, "random123(view, start, key)"},
    {"free", py_free, METH_VARARGS, "free(handle)"},
    {"op_slide", py_op_slide, METH_VARARGS, "op_slide(operand, rank, offset_change, shape_change, stride, shape, step_delay)"},
    {nullptr, nullptr, 0, nullptr},
};

PyModuleDef kModule = {PyModuleDef_HEAD_INIT, "_bhc", "fast bhc bridge calls", -1, kMethods,
                       nullptr, nullptr, nullptr, nullptr};

} // namespace

PyMODINIT_FUNC PyInit__bhc(void) { return PyModule_Create(&kModule); }
