"""ctypes mirror of include/sd2d.h and the marshalling of a mesh.Case into
the sd2d_mesh / sd2d_params structs.

Arrays handed to the C side are converted to the exact dtypes and
Fortran-contiguous layouts the ABI documents; the returned `keep` list
holds them alive for as long as the struct is used.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import mesh as M

SD2D_OK, SD2D_EINVAL, SD2D_EHIP, SD2D_ENOMEM, SD2D_ENONFINITE, \
    SD2D_EUNSUPPORTED = range(6)
SD2D_FP64, SD2D_MIXED = 0, 1
SD2D_TILE_DEFAULT, SD2D_TILE_STRIPS, SD2D_TILE_BLOCKS, SD2D_TILE_ROWPAIRS = 0, 1, 2, 3
# doubles of the static scratch of sd2d_compresid_noalloc (64 MiB)
SD2D_NOALLOC_MAX_DOUBLES = 8 * 1024 * 1024

_pi = C.POINTER(C.c_int32)
_pd = C.POINTER(C.c_double)


class Sd2dMesh(C.Structure):
    _fields_ = [("N", C.c_int32), ("ncell", C.c_int32), ("nface", C.c_int32),
                ("if2c", _pi),
                ("ninter", C.c_int32), ("ibfinter", _pi),
                ("ninlet", C.c_int32), ("ibfinl", _pi),
                ("noutlet", C.c_int32), ("ibfout", _pi),
                ("nwall", C.c_int32), ("ibfwal", _pi),
                ("nsymp", C.c_int32), ("ibfsymp", _pi),
                ("nfree", C.c_int32), ("ibfree", _pi),
                ("npair", C.c_int32), ("cyc2fpair", _pi),
                ("lmat", _pd), ("mmat", _pd), ("xs", _pd), ("xs_low", _pd),
                ("jac", _pd), ("s1", _pd), ("s2", _pd)]


class Sd2dParams(C.Structure):
    _fields_ = [("vismode", C.c_int32), ("artiv", C.c_int32),
                ("mu", C.c_double), ("eta0", C.c_double),
                ("alpha", C.c_double),
                ("rinf", C.c_double), ("pinf", C.c_double),
                ("uinf", C.c_double), ("vinf", C.c_double),
                ("bxinf", C.c_double), ("byinf", C.c_double),
                ("dxcycl", C.c_double), ("dycycl", C.c_double),
                ("ncellglob", C.c_int32), ("k_stage", C.c_int32),
                ("dt", C.c_double), ("precision", C.c_int32),
                ("tile_shape", C.c_int32)]


class Sd2dMaxima(C.Structure):
    _fields_ = [("eigvmax", C.c_double), ("umax", C.c_double),
                ("eigvmax_g", C.c_double), ("umax_g", C.c_double)]


class Sd2dDiag(C.Structure):
    _fields_ = [("resid_rms", C.c_double), ("divb_rms", C.c_double),
                ("eigvmax_g", C.c_double), ("umax_g", C.c_double),
                ("c_h", C.c_double), ("epsilon0", C.c_double),
                ("ctime", C.c_double), ("iter", C.c_int32),
                ("norms_valid", C.c_int32)]


def _iarr(a, keep):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel(order="F"))
    keep.append(a)
    return a.ctypes.data_as(_pi)


def _darr(a, keep):
    a = np.asarray(a, dtype=np.float64)
    a = np.ascontiguousarray(a.ravel(order="F"))
    keep.append(a)
    return a.ctypes.data_as(_pd)


def mesh_struct(case: M.Case):
    """(Sd2dMesh, keep) for a Case; the tables are exactly INIT_SETUP's."""
    keep = []
    t = case.topo
    L = t.lists
    s = Sd2dMesh()
    s.N = case.N
    s.ncell = case.ncell
    s.nface = t.if2c.shape[0]
    s.if2c = _iarr(np.asfortranarray(t.if2c), keep)
    s.ninter, s.ibfinter = len(t.ibfinter), _iarr(t.ibfinter, keep)
    s.ninlet, s.ibfinl = len(L["INLE"]), _iarr(L["INLE"], keep)
    s.noutlet, s.ibfout = len(L["OUTL"]), _iarr(L["OUTL"], keep)
    s.nwall, s.ibfwal = len(L["WALL"]), _iarr(L["WALL"], keep)
    s.nsymp, s.ibfsymp = len(L["SYMP"]), _iarr(L["SYMP"], keep)
    s.nfree, s.ibfree = len(L["FREE"]), _iarr(L["FREE"], keep)
    s.npair = t.cyc2fpair.shape[0]
    s.cyc2fpair = _iarr(np.asfortranarray(t.cyc2fpair), keep)
    s.lmat = _darr(case.lmat, keep)
    s.mmat = _darr(case.mmat, keep)
    s.xs = _darr(case.xs, keep)
    s.xs_low = _darr(case.xs_low if len(case.xs_low) else np.zeros(1), keep)
    s.jac = _darr(case.jac, keep)
    s.s1 = _darr(case.s1, keep)
    s.s2 = _darr(case.s2, keep)
    return s, keep


def params_struct(p: M.RunParams, ncellglob: int,
                  precision: int = SD2D_FP64,
                  tile_shape: int = SD2D_TILE_DEFAULT) -> Sd2dParams:
    s = Sd2dParams()
    s.vismode = p.vismode; s.artiv = p.artiv
    s.mu = p.mu; s.eta0 = p.eta0; s.alpha = p.alpha
    s.rinf = p.rinf; s.pinf = p.pinf; s.uinf = p.uinf; s.vinf = p.vinf
    s.bxinf = p.bxinf; s.byinf = p.byinf
    s.dxcycl = p.dxcycl; s.dycycl = p.dycycl
    s.ncellglob = ncellglob; s.k_stage = p.k_stage; s.dt = p.dt
    s.precision = precision
    s.tile_shape = tile_shape
    return s
