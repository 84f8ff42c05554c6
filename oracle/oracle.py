"""ctypes wrapper of the C restatement (liboracle_sd2d.so).

Test infrastructure: imported only by tests/, the smoke check and the CPU
baseline of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from sddc_amd import abi
from sddc_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle_sd2d.so")
# the same C source built with FMA contraction: a second legal IEEE
# evaluation of the reference's arithmetic, used only to measure the
# rounding floor of an input (tests/test_gpu_configs.py)
LIBPATH_FMA = os.path.join(HERE, "liboracle_sd2d_fma.so")
_libs = {}


def lib(fma: bool = False):
    path = LIBPATH_FMA if fma else LIBPATH
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(abi.Sd2dMesh), C.POINTER(abi.Sd2dParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_get_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_time.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_get_time.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_compresid.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_diag.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4
        L.orc_get_av.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_threads.restype = C.c_int
        _libs[path] = L
    return _libs[path]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """The reference's compresid / RK path on the CPU, same call contract
    as the device library."""

    def __init__(self, case: M.Case, fma: bool = False):
        self._L = lib(fma)
        self.case = case
        self.N, self.nc = case.N, case.ncell
        ms, self._keep = abi.mesh_struct(case)
        self._ps = abi.params_struct(case.params, case.ncell)
        self._ms = ms
        self.h = self._L.orc_create(C.byref(ms), C.byref(self._ps))
        if not self.h:
            raise RuntimeError("orc_create failed")

    def close(self):
        if self.h:
            self._L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, Q, maxima=None):
        Q = np.asfortranarray(Q, dtype=np.float64)
        mx = None
        if maxima is not None:
            mx = abi.Sd2dMaxima(*[float(x) for x in maxima])
        self._L.orc_set_state(self.h, _ptr(Q), C.byref(mx) if mx else None)

    def get_state(self):
        Q = np.zeros((9, self.N, self.N, self.nc), order="F")
        mx = abi.Sd2dMaxima()
        self._L.orc_get_state(self.h, _ptr(Q), C.byref(mx))
        return Q, np.array([mx.eigvmax, mx.umax, mx.eigvmax_g, mx.umax_g])

    def set_time(self, ctime, it=0):
        self._L.orc_set_time(self.h, float(ctime), int(it))

    def get_time(self):
        t = C.c_double(); i = C.c_int()
        self._L.orc_get_time(self.h, C.byref(t), C.byref(i))
        return t.value, i.value

    def compresid(self):
        r = np.zeros((9, self.N, self.N, self.nc), order="F")
        g = np.zeros((9, 2, self.N, self.N, self.nc), order="F")
        self._L.orc_compresid(self.h, _ptr(r), _ptr(g))
        return r, g

    def step(self, nsteps, norm_every=0):
        self._L.orc_step(self.h, int(nsteps), int(norm_every))

    def diag(self):
        v = [C.c_double() for _ in range(4)]
        self._L.orc_diag(self.h, *[C.byref(x) for x in v])
        return dict(zip(("c_h", "epsilon0", "resid_rms", "divb_rms"),
                        [x.value for x in v]))

    def get_av(self):
        s = np.zeros((2, self.nc), order="F"); m = np.zeros((2, self.nc), order="F")
        self._L.orc_get_av(self.h, _ptr(s), _ptr(m))
        return s, m


def threads() -> int:
    return lib().orc_threads()
