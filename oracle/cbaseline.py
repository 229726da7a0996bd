"""ctypes access to the oracle's C restatement (CPU baseline). TEST INFRASTRUCTURE ONLY."""
import ctypes
import os

import numpy as np

from .build import OUT

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(OUT):
            raise RuntimeError(f"{OUT} missing: run python -m oracle.build")
        lib = ctypes.CDLL(OUT)
        p, i64, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
        lib.nbody_oseen_f64.argtypes = [i64, p, i64, p, p, d, d, d, p]
        lib.nbody_stresslet_f64.argtypes = [i64, p, i64, p, p, p, d, d, d, p]
        _lib = lib
    return _lib


def _c(a):
    return np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1, 3))


def oseen(r_source, r_target, density, eta=1.0, reg=5e-3, eps=1e-5):
    s, t, f = _c(r_source), _c(r_target), _c(density)
    u = np.empty_like(t)
    load().nbody_oseen_f64(s.shape[0], s.ctypes.data, t.shape[0], t.ctypes.data, f.ctypes.data,
                           eta, reg, eps, u.ctypes.data)
    return u.ravel()


def stresslet(r_source, r_target, normal, density, eta=1.0, reg=5e-3, eps=1e-5):
    s, t, n, q = _c(r_source), _c(r_target), _c(normal), _c(density)
    u = np.empty_like(t)
    load().nbody_stresslet_f64(s.shape[0], s.ctypes.data, t.shape[0], t.ctypes.data,
                               n.ctypes.data, q.ctypes.data, eta, reg, eps, u.ctypes.data)
    return u.ravel()
