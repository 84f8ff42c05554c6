"""The C-ABI library on the CPU: it loads, exports every entry point that
include/sd2d.h declares, and rejects bad input with the documented error
codes before it touches a GPU (sd2d_create validates the mesh tables on
the host first).  No compute call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from sddc_amd import abi, mesh as M
from sddc_amd import solver

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "include", "sd2d.h")


def _lib():
    if not os.path.exists(solver.LIBPATH):
        pytest.fail(f"{solver.LIBPATH} is not built (run __graft_entry__.build())")
    return solver.load_library()


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sd2d_[a-z_]+)\s*\(", txt)))


def test_header_matches_python_exports():
    assert declared_functions() == sorted(solver.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    # and they are real exported text symbols of the shared object
    out = os.popen(f"nm -D --defined-only {solver.LIBPATH}").read()
    exported = set(re.findall(r"\sT\s(sd2d_\w+)", out))
    assert set(declared_functions()) <= exported


def test_struct_layouts():
    """The ctypes mirror has the C struct sizes (x86-64 SysV alignment)."""
    assert C.sizeof(abi.Sd2dMesh) == 3 * 4 + 4 + 8 + 7 * 16 + 7 * 8
    assert C.sizeof(abi.Sd2dParams) == 8 + 2 * 8 + 8 + 8 * 8 + 8 + 8 + 8
    assert C.sizeof(abi.Sd2dMaxima) == 32
    assert C.sizeof(abi.Sd2dDiag) == 7 * 8 + 8


def _small_case(N=4, **over):
    m = M.channel(4, 3, seed=1)
    p = M.RunParams(N=N, dxcycl=m.dxcycl, dycycl=m.dycycl, **over)
    return M.build_case(m, p)


def _create(case, mutate_mesh=None, mutate_params=None, precision=abi.SD2D_FP64):
    L = _lib()
    ms, keep = abi.mesh_struct(case)
    ps = abi.params_struct(case.params, case.ncell, precision)
    if mutate_mesh:
        mutate_mesh(ms, keep)
    if mutate_params:
        mutate_params(ps)
    h = C.c_void_p(1234)
    rc = L.sd2d_create(C.byref(ms), C.byref(ps), C.byref(h))
    msg = L.sd2d_last_error(None).decode()
    return rc, h, msg


def test_null_arguments():
    L = _lib()
    assert L.sd2d_create(None, None, None) == abi.SD2D_EINVAL
    h = C.c_void_p(77)
    assert L.sd2d_create(None, None, C.byref(h)) == abi.SD2D_EINVAL
    assert h.value is None                       # *out cleared on failure
    assert "NULL" in L.sd2d_last_error(None).decode()
    for f, args in [("sd2d_set_state", (None, None, None)),
                    ("sd2d_get_state", (None, None, None)),
                    ("sd2d_set_time", (None, 0.0, 0)),
                    ("sd2d_compresid", (None, None, None)),
                    ("sd2d_compresid_q", (None, None, None, None)),
                    ("sd2d_step", (None, 1, 0)),
                    ("sd2d_get_diag", (None, None)),
                    ("sd2d_get_av", (None, None, None, None)),
                    ("sd2d_write_restart", (None, b"x", None)),
                    ("sd2d_read_restart", (None, b"x")),
                    ("sd2d_set_profiling", (None, 1)),
                    ("sd2d_get_profile", (None, None, None))]:
        assert getattr(L, f)(*args) == abi.SD2D_EINVAL, f
    L.sd2d_destroy(None)                          # a no-op


@pytest.mark.parametrize("N", [2, 6])
def test_unsupported_order(N):
    case = _small_case(4)

    def mut(ms, keep):
        ms.N = N
    rc, h, msg = _create(case, mut)
    assert rc == abi.SD2D_EUNSUPPORTED and "N must be" in msg
    assert h.value is None


def test_bad_k_stage():
    rc, _, msg = _create(_small_case(), mutate_params=lambda p: setattr(p, "k_stage", 4))
    assert rc == abi.SD2D_EINVAL and "k_stage" in msg


def test_viscous_mixed_rejected():
    """The viscous/resistive path (vismode=1, mu/eta0 != 0) runs in FP64
    only: asking for it in mixed precision is refused rather than silently
    run in FP64."""
    rc, _, msg = _create(_small_case(vismode=1, mu=0.01), precision=abi.SD2D_MIXED)
    assert rc == abi.SD2D_EUNSUPPORTED and "viscous" in msg


@pytest.mark.parametrize("vismode", [0, 2])
def test_vismode_other_than_1_rejected(vismode):
    """vismode 0 has no defined result in the reference (its Powell source
    and divB norm read a nablaQs that is never computed, compflux.f90:13-15
    against compresid.f90:58), so it is refused, as is any value but 1."""
    rc, h, msg = _create(_small_case(), mutate_params=lambda p: setattr(p, "vismode", vismode))
    assert rc == abi.SD2D_EUNSUPPORTED and "vismode" in msg
    assert h.value is None
    L = _lib()
    case = _small_case()
    ms, keep = abi.mesh_struct(case)
    ps = abi.params_struct(case.params, case.ncell)
    ps.vismode = vismode
    meshes = (abi.Sd2dMesh * 1)(ms)
    params = (abi.Sd2dParams * 1)(ps)
    e = C.c_void_p(99)
    rc = L.sd2d_ens_create(1, meshes, params, C.byref(e))
    assert rc == abi.SD2D_EUNSUPPORTED and "vismode" in L.sd2d_ens_last_error(None).decode()
    assert e.value is None


def test_check_finite_arguments():
    L = _lib()
    assert L.sd2d_set_check_finite(None, 1) == abi.SD2D_EINVAL
    assert L.sd2d_ens_set_check_finite(None, 1) == abi.SD2D_EINVAL
    assert L.sd2d_ens_nonfinite(None, None) == abi.SD2D_EINVAL


def test_unknown_precision():
    rc, _, msg = _create(_small_case(), precision=7)
    assert rc == abi.SD2D_EINVAL and "precision" in msg


def test_incomplete_mesh():
    def mut(ms, keep):
        ms.jac = C.POINTER(C.c_double)()
    rc, _, msg = _create(_small_case(), mut)
    assert rc == abi.SD2D_EINVAL and "incomplete" in msg


def test_face_list_out_of_range():
    case = _small_case()

    def mut(ms, keep):
        a = np.array(case.topo.ibfinter, dtype=np.int32)
        a[0] = case.topo.if2c.shape[0] + 5
        keep.append(a)
        ms.ibfinter = a.ctypes.data_as(C.POINTER(C.c_int32))
    rc, _, msg = _create(case, mut)
    assert rc == abi.SD2D_EINVAL and "IBFINTER" in msg


def test_unlisted_face_is_a_processor_interface():
    """A face that is in no list (on a multi-rank mesh: a processor
    interface) is refused: this library runs one rank's whole mesh."""
    case = _small_case()

    def mut(ms, keep):
        ms.ninter = ms.ninter - 1
    rc, _, msg = _create(case, mut)
    assert rc == abi.SD2D_EINVAL and "interface" in msg


def _perturbed_s1(case, rel):
    def mut(ms, keep):
        s1 = np.array(case.s1, dtype=np.float64, order="F").copy(order="F")
        s1[0, 0, 0, 1, 0] *= 1.0 + rel      # one face-point normal component of cell 1
        a = np.ascontiguousarray(s1.ravel(order="F"))
        keep.append(a)
        ms.s1 = a.ctypes.data_as(C.POINTER(C.c_double))
    return mut


def test_nonconforming_normals_refused():
    """The two cells of a face must see one normal up to sign: normals that
    differ by more than rounding (1e-9 of their length) are refused, as a
    non-conforming face."""
    case = _small_case()
    rc, _, msg = _create(case, _perturbed_s1(case, 1e-6))
    assert rc == abi.SD2D_EUNSUPPORTED and "non-conforming" in msg, msg


def test_rounding_different_normals_accepted():
    """Normals that differ by rounding (as the two cells of a curved 8-node
    face form them) pass the face-table check: the right cell then takes
    the left cell's normal.  Without a GPU the call stops later, at the
    device, not at the check."""
    case = _small_case()
    rc, ctx, msg = _create(case, _perturbed_s1(case, 4e-16))
    if rc == abi.SD2D_OK:
        solver._lib.sd2d_destroy(ctx)
    assert "non-conforming" not in msg and "normal" not in msg, msg


def test_product_path_needs_the_library(tmp_path):
    """No CPU fallback: a missing library raises."""
    with pytest.raises(solver.Sd2dError):
        solver._lib_saved = solver._lib
        try:
            solver._lib = None
            solver.load_library(str(tmp_path / "missing.so"))
        finally:
            solver._lib = solver._lib_saved


def test_fortran_binding_layout(tmp_path):
    """The bind(C) types of sddc_amd/fortran/sd2d_c.f90 have the sizes of
    the C structs (compiled with flang; skipped where flang is absent)."""
    import shutil
    import subprocess
    fc = shutil.which("flang") or ("/opt/rocm/bin/amdflang"
                                   if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        pytest.skip("no Fortran compiler")
    src = os.path.join(os.path.dirname(HERE), "sddc_amd", "fortran", "sd2d_c.f90")
    prog = tmp_path / "sizes.f90"
    prog.write_text(
        "program sizes\n use iso_c_binding\n use sd2d_c\n"
        " type(sd2d_mesh_t) :: m\n type(sd2d_params_t) :: p\n"
        " type(sd2d_maxima_t) :: x\n type(sd2d_diag_t) :: d\n"
        " print *, c_sizeof(m), c_sizeof(p), c_sizeof(x), c_sizeof(d)\n"
        "end program sizes\n")
    exe = tmp_path / "sizes"
    libdir = os.path.dirname(_lib()._name)
    r = subprocess.run([fc, "-module-dir", str(tmp_path), src, str(prog), "-o", str(exe),
                        f"-L{libdir}", "-lsd2d", f"-Wl,-rpath,{libdir}",
                        "-Wl,-rpath,/opt/rocm/lib"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True,
                                          text=True).stdout.split()]
    assert got == [C.sizeof(abi.Sd2dMesh), C.sizeof(abi.Sd2dParams),
                   C.sizeof(abi.Sd2dMaxima), C.sizeof(abi.Sd2dDiag)]


def _ens_create(cases, mutate=None):
    L = _lib()
    n = len(cases)
    meshes = (abi.Sd2dMesh * n)(); params = (abi.Sd2dParams * n)(); keep = []
    for i, c in enumerate(cases):
        ms, k = abi.mesh_struct(c)
        meshes[i] = ms; keep.append(k)
        params[i] = abi.params_struct(c.params, c.ncell)
    if mutate:
        mutate(meshes, params)
    h = C.c_void_p(1234)
    rc = L.sd2d_ens_create(n, meshes, params, C.byref(h))
    return rc, h, L.sd2d_ens_last_error(None).decode()


def test_ensemble_rejects_bad_input():
    """Ensemble creation refuses, before touching the device: no members,
    FiveStageRk, the viscous path, mixed precision settings."""
    L = _lib()
    h = C.c_void_p(1234)
    assert L.sd2d_ens_create(0, None, None, C.byref(h)) == abi.SD2D_EINVAL
    cs = [_small_case(), _small_case(N=3)]
    rc, h, msg = _ens_create(cs, lambda m, p: setattr(p[1], "k_stage", 5))
    assert rc == abi.SD2D_EUNSUPPORTED and "RK_S5P3" in msg and h.value is None
    rc, _, msg = _ens_create(cs, lambda m, p: (setattr(p[0], "mu", 0.01)))
    assert rc == abi.SD2D_EUNSUPPORTED and "inviscid" in msg
    rc, _, msg = _ens_create(cs, lambda m, p: setattr(p[1], "precision", abi.SD2D_MIXED))
    assert rc == abi.SD2D_EINVAL and "precision" in msg
    rc, _, msg = _ens_create(cs, lambda m, p: setattr(m[0], "N", 7))
    assert rc == abi.SD2D_EUNSUPPORTED and "N must be" in msg


def test_fortran_module_declares_every_entry_point():
    """sddc_amd/fortran/sd2d_c.f90 binds every function include/sd2d.h
    declares, under its C name."""
    import re
    src = open(os.path.join(os.path.dirname(HERE), "sddc_amd", "fortran", "sd2d_c.f90")).read()
    bound = sorted(set(re.findall(r'bind\(C,\s*name="(sd2d_\w+)"\)', src)))
    assert bound == declared_functions()
    # and the module (private by default) makes every one of them public
    m = re.search(r"^\s*public ::\s*(sd2d_.*?)\n\s*\n", src, re.S | re.M)
    assert m, "no public statement of the entry points"
    public = set(re.findall(r"\b(sd2d_\w+)\b", m.group(1)))
    assert set(bound) <= public, sorted(set(bound) - public)


def test_noalloc_limit_matches_header():
    """The scratch limit of sd2d_compresid_noalloc in the ctypes mirror is
    the header's SD2D_NOALLOC_MAX_DOUBLES."""
    m = re.search(r"#define\s+SD2D_NOALLOC_MAX_DOUBLES\s+\(([\d\s\*]+)\)", open(HEADER).read())
    assert m, "SD2D_NOALLOC_MAX_DOUBLES not defined in include/sd2d.h"
    val = 1
    for f in m.group(1).split("*"):
        val *= int(f)
    assert val == abi.SD2D_NOALLOC_MAX_DOUBLES
