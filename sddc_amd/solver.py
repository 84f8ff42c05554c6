"""Python host over the C ABI of libsd2d.so (include/sd2d.h).

Mirrors the reference's interface for this path: `compresid()` has the
per-call contract of `subroutine compresid` (compresid.f90:1-80), `step()`
advances like RK_S5P3 / FiveStageRk called from `iterations`
(iterations.f90:19-31), and the state and running maxima move in and out in
the reference's Fortran layouts.  There is no fallback: if the device
library is missing or fails, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from . import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
# the library built by sddc_amd.build; SD2D_LIB names another build of the
# same ABI (e.g. a variant under comparison)
LIBPATH = os.environ.get("SD2D_LIB") or os.path.join(HERE, "lib", "libsd2d.so")
_lib = None

EXPORTS = ("sd2d_create", "sd2d_destroy", "sd2d_last_error", "sd2d_set_state",
           "sd2d_get_state", "sd2d_set_time", "sd2d_compresid",
           "sd2d_compresid_q", "sd2d_compresid_noalloc", "sd2d_step", "sd2d_rk_stage", "sd2d_get_diag", "sd2d_get_av", "sd2d_plot_nodes",
           "sd2d_write_restart", "sd2d_read_restart", "sd2d_write_checkpoint",
           "sd2d_read_checkpoint", "sd2d_ens_write_checkpoint", "sd2d_ens_read_checkpoint",
           "sd2d_set_profiling", "sd2d_get_profile",
           "sd2d_ens_create", "sd2d_ens_destroy", "sd2d_ens_last_error",
           "sd2d_ens_set_state", "sd2d_ens_get_state", "sd2d_ens_set_time",
           "sd2d_ens_get_time", "sd2d_ens_step", "sd2d_ens_tune",
           "sd2d_set_check_finite", "sd2d_ens_set_check_finite", "sd2d_ens_nonfinite",
           "sd2d_fwd", "sd2d_vjp", "sd2d_compresid_vjp", "sd2d_adj_save", "sd2d_adj_load")


class Sd2dError(RuntimeError):
    pass


class Sd2dNonFinite(Sd2dError):
    """The opt-in finiteness check found a NaN or an infinity."""


def load_library(path: str = LIBPATH):
    """Load libsd2d.so and declare the ABI; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise Sd2dError(f"{path} is not built; run `python -m sddc_amd.build`")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.sd2d_create.argtypes = [C.POINTER(abi.Sd2dMesh), C.POINTER(abi.Sd2dParams),
                              C.POINTER(vp)]
    L.sd2d_destroy.argtypes = [vp]
    L.sd2d_destroy.restype = None
    L.sd2d_last_error.argtypes = [vp]
    L.sd2d_last_error.restype = C.c_char_p
    L.sd2d_set_state.argtypes = [vp, vp, C.POINTER(abi.Sd2dMaxima)]
    L.sd2d_get_state.argtypes = [vp, vp, C.POINTER(abi.Sd2dMaxima)]
    L.sd2d_set_time.argtypes = [vp, C.c_double, C.c_int32]
    L.sd2d_compresid.argtypes = [vp, vp, vp]
    L.sd2d_compresid_q.argtypes = [vp, vp, vp, vp]
    L.sd2d_compresid_noalloc.argtypes = [vp, vp, vp]
    L.sd2d_step.argtypes = [vp, C.c_int32, C.c_int32]
    L.sd2d_rk_stage.argtypes = [vp, C.c_int32]
    L.sd2d_get_diag.argtypes = [vp, C.POINTER(abi.Sd2dDiag)]
    L.sd2d_get_av.argtypes = [vp, vp, vp, C.POINTER(C.c_double)]
    L.sd2d_write_restart.argtypes = [vp, C.c_char_p, vp]
    L.sd2d_read_restart.argtypes = [vp, C.c_char_p]
    L.sd2d_write_checkpoint.argtypes = [vp, C.c_char_p, vp]
    L.sd2d_read_checkpoint.argtypes = [vp, C.c_char_p]
    L.sd2d_ens_write_checkpoint.argtypes = [vp, C.c_int32, C.c_char_p, vp]
    L.sd2d_ens_read_checkpoint.argtypes = [vp, C.c_int32, C.c_char_p]
    L.sd2d_set_profiling.argtypes = [vp, C.c_int32]
    L.sd2d_plot_nodes.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp]
    L.sd2d_get_profile.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    L.sd2d_ens_create.argtypes = [C.c_int32, C.POINTER(abi.Sd2dMesh),
                                  C.POINTER(abi.Sd2dParams), C.POINTER(vp)]
    L.sd2d_ens_destroy.argtypes = [vp]
    L.sd2d_ens_destroy.restype = None
    L.sd2d_ens_last_error.argtypes = [vp]
    L.sd2d_ens_last_error.restype = C.c_char_p
    L.sd2d_ens_set_state.argtypes = [vp, C.c_int32, vp, C.POINTER(abi.Sd2dMaxima)]
    L.sd2d_ens_get_state.argtypes = [vp, C.c_int32, vp, C.POINTER(abi.Sd2dMaxima)]
    L.sd2d_ens_set_time.argtypes = [vp, C.c_int32, C.c_double, C.c_int32]
    L.sd2d_ens_get_time.argtypes = [vp, C.c_int32, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int32)]
    L.sd2d_ens_step.argtypes = [vp, C.c_int32]
    L.sd2d_ens_tune.argtypes = [vp, C.c_int32, vp]
    L.sd2d_set_check_finite.argtypes = [vp, C.c_int32]
    L.sd2d_ens_set_check_finite.argtypes = [vp, C.c_int32]
    L.sd2d_ens_nonfinite.argtypes = [vp, vp]
    L.sd2d_fwd.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    L.sd2d_vjp.argtypes = [vp, vp, C.c_double, vp, vp, vp]
    L.sd2d_compresid_vjp.argtypes = [vp, vp, vp, C.POINTER(C.c_double), vp]
    L.sd2d_adj_save.argtypes = [vp, C.c_char_p]
    L.sd2d_adj_load.argtypes = [vp, C.c_char_p]
    for f in EXPORTS:
        if f not in ("sd2d_destroy", "sd2d_last_error", "sd2d_ens_destroy",
                     "sd2d_ens_last_error"):
            getattr(L, f).restype = C.c_int
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Sd2d:
    """One device context for one mesh (one GPU)."""

    def __init__(self, case: M.Case, precision: int = abi.SD2D_FP64,
                 tile_shape: int = abi.SD2D_TILE_DEFAULT):
        self.L = load_library()
        self.case = case
        self.N, self.nc = case.N, case.ncell
        ms, self._keep = abi.mesh_struct(case)
        ps = abi.params_struct(case.params, case.ncell, precision, tile_shape)
        h = C.c_void_p()
        rc = self.L.sd2d_create(C.byref(ms), C.byref(ps), C.byref(h))
        if rc != 0:
            raise Sd2dError(f"sd2d_create: {rc}: {self.L.sd2d_last_error(None).decode()}")
        self.h = h
        self._keep = None   # tables were copied by sd2d_create

    def _check(self, rc, what):
        if rc != 0:
            cls = Sd2dNonFinite if rc == abi.SD2D_ENONFINITE else Sd2dError
            raise cls(f"{what}: {rc}: {self.L.sd2d_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.sd2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------
    def set_state(self, Q, maxima=None):
        Q = np.asfortranarray(Q, dtype=np.float64)
        assert Q.shape == (9, self.N, self.N, self.nc)
        mx = abi.Sd2dMaxima(*[float(x) for x in maxima]) if maxima is not None else None
        self._check(self.L.sd2d_set_state(self.h, _ptr(Q), C.byref(mx) if mx else None),
                    "sd2d_set_state")

    def get_state(self):
        Q = np.zeros((9, self.N, self.N, self.nc), order="F")
        mx = abi.Sd2dMaxima()
        self._check(self.L.sd2d_get_state(self.h, _ptr(Q), C.byref(mx)), "sd2d_get_state")
        return Q, np.array([mx.eigvmax, mx.umax, mx.eigvmax_g, mx.umax_g])

    def set_time(self, ctime, it=0):
        self._check(self.L.sd2d_set_time(self.h, float(ctime), int(it)), "sd2d_set_time")

    def maxima(self):
        """The running maxima alone [eigvmax, umax, eigvmax_g, umax_g]
        (allowed between the stages of a step)."""
        mx = abi.Sd2dMaxima()
        self._check(self.L.sd2d_get_state(self.h, None, C.byref(mx)), "sd2d_get_state")
        return np.array([mx.eigvmax, mx.umax, mx.eigvmax_g, mx.umax_g])

    # -- the path ----------------------------------------------------------
    def compresid(self, want_resid=True, want_nabla=True):
        r = np.zeros((9, self.N, self.N, self.nc), order="F") if want_resid else None
        g = np.zeros((9, 2, self.N, self.N, self.nc), order="F") if want_nabla else None
        self._check(self.L.sd2d_compresid(self.h, _ptr(r), _ptr(g)), "sd2d_compresid")
        return r, g

    def step(self, nsteps, norm_every=0):
        self._check(self.L.sd2d_step(self.h, int(nsteps), int(norm_every)), "sd2d_step")

    def rk_stage(self, stage):
        """One stage (1..5) of the current step, in order."""
        self._check(self.L.sd2d_rk_stage(self.h, int(stage)), "sd2d_rk_stage")

    def diag(self):
        d = abi.Sd2dDiag()
        self._check(self.L.sd2d_get_diag(self.h, C.byref(d)), "sd2d_get_diag")
        return {k: getattr(d, k) for k, _ in abi.Sd2dDiag._fields_}

    def plot_nodes(self, topo, muav=None):
        """Nodal values of the Tecplot output on the plot mesh `topo`
        (sddc_amd.tecplot.plot_topology): plotq (9, nnode) and, with the
        per-cell muAV (2, NCELL) given, plotav (2, nnode)."""
        nn = topo.nnode
        xadj = np.ascontiguousarray(topo.xadj, dtype=np.int32)
        occ = np.ascontiguousarray(topo.contrib, dtype=np.int32)
        q = np.zeros((9, nn), order="F")
        av = np.zeros((2, nn), order="F") if muav is not None else None
        mu = None if muav is None else np.asfortranarray(muav, dtype=np.float64)
        self._check(self.L.sd2d_plot_nodes(self.h, int(topo.nre), nn, _ptr(xadj), _ptr(occ),
                                           _ptr(mu), _ptr(q), _ptr(av)), "sd2d_plot_nodes")
        return q, av

    def write_restart(self, path, icg=None):
        """Reference-format restart file of the device state."""
        g = None if icg is None else np.ascontiguousarray(icg, dtype=np.int32)
        self._check(self.L.sd2d_write_restart(self.h, str(path).encode(), _ptr(g)),
                    "sd2d_write_restart")

    def read_restart(self, path):
        """Load a reference-format restart file (maxima reset to 0)."""
        self._check(self.L.sd2d_read_restart(self.h, str(path).encode()), "sd2d_read_restart")

    def write_checkpoint(self, path, icg=None):
        """Reference-format restart file plus the running-maxima sidecar
        <path>.max, from which a run resumes bit for bit."""
        g = None if icg is None else np.ascontiguousarray(icg, dtype=np.int32)
        self._check(self.L.sd2d_write_checkpoint(self.h, str(path).encode(), _ptr(g)),
                    "sd2d_write_checkpoint")

    def read_checkpoint(self, path):
        """Load a checkpoint: state, iter, ctime and the running maxima."""
        self._check(self.L.sd2d_read_checkpoint(self.h, str(path).encode()),
                    "sd2d_read_checkpoint")

    def set_check_finite(self, every: int):
        """Opt-in NaN/Inf check of the state every `every` steps (0: off);
        a hit makes step() raise Sd2dNonFinite."""
        self._check(self.L.sd2d_set_check_finite(self.h, int(every)), "sd2d_set_check_finite")

    def set_profiling(self, on: bool):
        self._check(self.L.sd2d_set_profiling(self.h, int(bool(on))), "sd2d_set_profiling")

    def get_profile(self):
        """(summed stage-kernel milliseconds, number of stage launches)."""
        ms = C.c_double(); n = C.c_int32()
        self._check(self.L.sd2d_get_profile(self.h, C.byref(ms), C.byref(n)), "sd2d_get_profile")
        return ms.value, n.value

    # -- adjoint -------------------------------------------------------------
    def fwd(self, nsteps, ckpt=20):
        """nsteps RK_S5P3 steps with checkpoints; returns the time-averaged
        wall drag J."""
        J = C.c_double()
        self._check(self.L.sd2d_fwd(self.h, int(nsteps), int(ckpt), C.byref(J)), "sd2d_fwd")
        return J.value

    def vjp(self, qbar_final=None, jbar=1.0, want_chbar=False):
        """Cotangents of the initial state and of the free stream
        (rinf, pinf, uinf, vinf, bxinf, byinf) after fwd(); optionally the
        per-evaluation c_h cotangents."""
        qf = None if qbar_final is None else np.asfortranarray(qbar_final, dtype=np.float64)
        q0 = np.zeros((9, self.N, self.N, self.nc), order="F")
        pb = np.zeros(6)
        ch = np.zeros(5 * 100000) if want_chbar else None
        self._check(self.L.sd2d_vjp(self.h, _ptr(qf), float(jbar), _ptr(q0), _ptr(pb), _ptr(ch)),
                    "sd2d_vjp")
        return q0, pb, ch

    def adj_save(self, prefix):
        """The checkpoints of the last fwd() as <prefix>.NNNN.dat restart
        files (with maxima sidecars) and the manifest <prefix>.adj."""
        self._check(self.L.sd2d_adj_save(self.h, str(prefix).encode()), "sd2d_adj_save")

    def adj_load(self, prefix):
        """Load what adj_save wrote: the context is then as after fwd(),
        ready for vjp()."""
        self._check(self.L.sd2d_adj_load(self.h, str(prefix).encode()), "sd2d_adj_load")

    def compresid_vjp(self, rbar):
        rb = np.asfortranarray(rbar, dtype=np.float64)
        q = np.zeros((9, self.N, self.N, self.nc), order="F")
        ch = C.c_double()
        pb = np.zeros(6)
        self._check(self.L.sd2d_compresid_vjp(self.h, _ptr(rb), _ptr(q), C.byref(ch), _ptr(pb)),
                    "sd2d_compresid_vjp")
        return q, ch.value, pb

    def get_av(self):
        s = np.zeros((2, self.nc), order="F"); m = np.zeros((2, self.nc), order="F")
        e = C.c_double()
        self._check(self.L.sd2d_get_av(self.h, _ptr(s), _ptr(m), C.byref(e)), "sd2d_get_av")
        return s, m, e.value


class Sd2dEnsemble:
    """Many independent runs (meshes, orders, free streams, time steps)
    stepped together on one GPU through the sd2d_ens_* entry points."""

    def __init__(self, cases, precision: int = abi.SD2D_FP64,
                 tile_shape: int = abi.SD2D_TILE_DEFAULT):
        self.L = load_library()
        self.cases = list(cases)
        n = len(self.cases)
        meshes = (abi.Sd2dMesh * n)()
        params = (abi.Sd2dParams * n)()
        self._keep = []   # the arrays stay alive: sd2d_ens_tune reads them again
        for i, c in enumerate(self.cases):
            ms, keep = abi.mesh_struct(c)
            meshes[i] = ms
            params[i] = abi.params_struct(c.params, c.ncell, precision, tile_shape)
            self._keep.append(keep)
        self._structs = (meshes, params)
        h = C.c_void_p()
        rc = self.L.sd2d_ens_create(n, meshes, params, C.byref(h))
        if rc != 0:
            raise Sd2dError(f"sd2d_ens_create: {rc}: {self.L.sd2d_ens_last_error(None).decode()}")
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            cls = Sd2dNonFinite if rc == abi.SD2D_ENONFINITE else Sd2dError
            raise cls(f"{what}: {rc}: {self.L.sd2d_ens_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.sd2d_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, i, Q, maxima=None):
        c = self.cases[i]
        Q = np.asfortranarray(Q, dtype=np.float64)
        assert Q.shape == (9, c.N, c.N, c.ncell)
        mx = abi.Sd2dMaxima(*[float(x) for x in maxima]) if maxima is not None else None
        self._check(self.L.sd2d_ens_set_state(self.h, i, _ptr(Q), C.byref(mx) if mx else None),
                    "sd2d_ens_set_state")

    def get_state(self, i):
        c = self.cases[i]
        Q = np.zeros((9, c.N, c.N, c.ncell), order="F")
        mx = abi.Sd2dMaxima()
        self._check(self.L.sd2d_ens_get_state(self.h, i, _ptr(Q), C.byref(mx)),
                    "sd2d_ens_get_state")
        return Q, np.array([mx.eigvmax, mx.umax, mx.eigvmax_g, mx.umax_g])

    def get_time(self, i):
        t = C.c_double(); it = C.c_int32()
        self._check(self.L.sd2d_ens_get_time(self.h, i, C.byref(t), C.byref(it)),
                    "sd2d_ens_get_time")
        return t.value, it.value

    def write_checkpoint(self, i, path, icg=None):
        """Member i's checkpoint (restart file plus maxima sidecar)."""
        g = None if icg is None else np.ascontiguousarray(icg, dtype=np.int32)
        self._check(self.L.sd2d_ens_write_checkpoint(self.h, int(i), str(path).encode(), _ptr(g)),
                    "sd2d_ens_write_checkpoint")

    def read_checkpoint(self, i, path):
        """Load a checkpoint into member i (state, maxima, iter, ctime)."""
        self._check(self.L.sd2d_ens_read_checkpoint(self.h, int(i), str(path).encode()),
                    "sd2d_ens_read_checkpoint")

    def step(self, nsteps):
        self._check(self.L.sd2d_ens_step(self.h, int(nsteps)), "sd2d_ens_step")

    def set_check_finite(self, every: int):
        """Opt-in NaN/Inf check of every member every `every` steps (0: off);
        a hit makes step() raise Sd2dNonFinite, nonfinite() names them."""
        self._check(self.L.sd2d_ens_set_check_finite(self.h, int(every)),
                    "sd2d_ens_set_check_finite")

    def nonfinite(self):
        """Per member: 1 when the latest check found a NaN or an infinity."""
        f = np.zeros(len(self.cases), dtype=np.int32)
        self._check(self.L.sd2d_ens_nonfinite(self.h, _ptr(f)), "sd2d_ens_nonfinite")
        return f

    def tune(self, trial_steps=3):
        """Autotune the tile shape of every order present; returns
        {N: shape}."""
        sh = np.full(3, -1, dtype=np.int32)
        self._check(self.L.sd2d_ens_tune(self.h, int(trial_steps), _ptr(sh)), "sd2d_ens_tune")
        return {n + 3: int(v) for n, v in enumerate(sh) if v >= 0}
