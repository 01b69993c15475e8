"""Batched execution of many AXPY / dot / GEMV instances of mixed precision in
one hipGraph launch (mpgfx_batch_*, include/mpgfx.h).

    tasks = [batch.axpy(a, x, y, out), batch.dot(x, y, out), batch.gemv(A, x, m, n, out)]
    b = batch.Batch(tasks, tuning=True, cache="tune.json")
    b.run()            # asynchronous on the current torch stream

Dot and GEMV use the reference's per-operation k-order fold
(tools/linalg_fast.cu:37-51).  The Batch keeps references to the arrays it was
built from; they must stay alive (and keep their addresses) while it is used.
"""
import ctypes

from . import _lib
from .arrays import MPArray

AXPY, DOT, GEMV = 0, 1, 2


class Task(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pb", ctypes.c_int32), ("cplx", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("m", ctypes.c_int64), ("n", ctypes.c_int64),
                ("in0_limbs", ctypes.c_void_p), ("in0_se", ctypes.c_void_p),
                ("in1_limbs", ctypes.c_void_p), ("in1_se", ctypes.c_void_p),
                ("in2_limbs", ctypes.c_void_p), ("in2_se", ctypes.c_void_p),
                ("out_limbs", ctypes.c_void_p), ("out_se", ctypes.c_void_p)]


assert ctypes.sizeof(Task) == 96


def _ptrs(x):
    return x.ptrs() if x is not None else (None, None)


def _check(what, pb, cplx, *arrs):
    """every array of a task: a device MP array of the task's precision and
    kind, with its expected length (a mismatch would be an out-of-bounds
    access on the device)"""
    for x, n in arrs:
        if not isinstance(x, MPArray) or not x.on_device:
            raise TypeError(f"{what}: expected device MPArrays")
        if x.pb != pb or x.cplx != cplx:
            raise ValueError(f"{what}: precision / kind mismatch (pb={x.pb} cplx={x.cplx}, want pb={pb} cplx={cplx})")
        if x.n != n:
            raise ValueError(f"{what}: array of {x.n} values where {n} are needed")


def axpy(a, x, y, out):
    """out = a*x + y; out may be y itself (in place)"""
    _check("batch.axpy", x.pb, x.cplx, (a, 1), (x, x.n), (y, x.n), (out, x.n))
    t = Task(AXPY, x.pb, int(x.cplx), 0, 1, x.n, *_ptrs(a), *_ptrs(x), *_ptrs(y), *_ptrs(out))
    return t, (a, x, y, out)


def dot(x, y, out):
    _check("batch.dot", x.pb, x.cplx, (x, x.n), (y, x.n), (out, 1))
    t = Task(DOT, x.pb, int(x.cplx), 0, 1, x.n, *_ptrs(x), *_ptrs(y), None, None, *_ptrs(out))
    return t, (x, y, out)


def gemv(A, x, m, n, out):
    _check("batch.gemv", x.pb, x.cplx, (A, m * n), (x, n), (out, m))
    t = Task(GEMV, x.pb, int(x.cplx), 0, m, n, *_ptrs(A), *_ptrs(x), None, None, *_ptrs(out))
    return t, (A, x, out)


class Batch:
    def __init__(self, tasks, tuning=False, cache=None):
        import torch
        self._keep = [k for _, k in tasks]
        arr = (Task * max(1, len(tasks)))(*[t for t, _ in tasks])
        h = ctypes.c_void_p(0)
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        c = cache.encode() if cache else None
        _lib.check("mpgfx_batch_create",
                   _lib.lib().mpgfx_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(tasks), int(bool(tuning)),
                                                 c, s, ctypes.byref(h)))
        self._h = h

    def run(self):
        import torch
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check("mpgfx_batch_run", _lib.lib().mpgfx_batch_run(self._h, s))

    def info(self):
        b, c = ctypes.c_size_t(0), ctypes.c_int(0)
        _lib.check("mpgfx_batch_info", _lib.lib().mpgfx_batch_info(self._h, ctypes.byref(b), ctypes.byref(c)))
        return {"scratch_bytes": b.value, "classes": c.value}

    def close(self):
        if self._h:
            _lib.lib().mpgfx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
