/* CPython fast path for Simulator.amplitudes: basis-index arrays and the
 * result array are passed through the buffer protocol straight to the C ABI
 * (pq_sim_amplitudes), with the GIL released during the call.  ctypes spends
 * about 2.5 us per call converting the same arguments; this path spends a
 * fraction of that.  The Python wrapper takes it for C-contiguous uint64
 * arrays and keeps the ctypes call for everything else. */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <stdint.h>
#include "polyq_amd.h"

/* amplitudes(handle: int, xs, ys, out, form: int) -> int (the C ABI status).
 * xs, ys: n uint64 each; out: n complex128 (2n doubles), writable. */
static PyObject* fast_amplitudes(PyObject* self, PyObject* args) {
    unsigned long long h;
    PyObject *xo, *yo, *oo;
    int form;
    (void)self;
    if (!PyArg_ParseTuple(args, "KOOOi", &h, &xo, &yo, &oo, &form)) return NULL;
    Py_buffer bx, by, bo;
    if (PyObject_GetBuffer(xo, &bx, PyBUF_C_CONTIGUOUS) != 0) return NULL;
    if (PyObject_GetBuffer(yo, &by, PyBUF_C_CONTIGUOUS) != 0) { PyBuffer_Release(&bx); return NULL; }
    if (PyObject_GetBuffer(oo, &bo, PyBUF_C_CONTIGUOUS | PyBUF_WRITABLE) != 0) { PyBuffer_Release(&bx); PyBuffer_Release(&by); return NULL; }
    const Py_ssize_t n = bx.len / 8;
    if (h == 0 || bx.len % 8 != 0 || by.len != bx.len || bo.len != 16 * n) {
        PyBuffer_Release(&bx); PyBuffer_Release(&by); PyBuffer_Release(&bo);
        PyErr_SetString(PyExc_ValueError, "amplitudes: xs and ys need n uint64 values and out 2n doubles");
        return NULL;
    }
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = pq_sim_amplitudes((pq_sim*)(uintptr_t)h, (size_t)n, (const uint64_t*)bx.buf, (const uint64_t*)by.buf, form, (double*)bo.buf);
    Py_END_ALLOW_THREADS
    PyBuffer_Release(&bx); PyBuffer_Release(&by); PyBuffer_Release(&bo);
    return PyLong_FromLong(rc);
}

static PyMethodDef methods[] = {
    {"amplitudes", fast_amplitudes, METH_VARARGS, "amplitudes(handle, xs, ys, out, form) -> status"},
    {NULL, NULL, 0, NULL},
};
static struct PyModuleDef module = {PyModuleDef_HEAD_INIT, "_pqfast", NULL, -1, methods, NULL, NULL, NULL, NULL};
PyMODINIT_FUNC PyInit__pqfast(void) { return PyModule_Create(&module); }
