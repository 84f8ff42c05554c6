"""sd2d_compresid_noalloc, the allocation-free form of sd2d_compresid, called
through the C ABI: on two meshes in a row (the second smaller, so the
static scratch still holds the first call's larger contents) its resid,
nablaQs, running maxima and c_h equal those of sd2d_compresid on the same
state bit for bit, over two consecutive calls each (the second with the
lagged c_h of the first).  A mesh whose nablaQs exceeds the static buffer
is refused with SD2D_EINVAL before any device work."""
import ctypes as C

import numpy as np
import pytest

from sddc_amd import abi
from sddc_amd import mesh as M

pytestmark = pytest.mark.gpu


def _noalloc(g, want_nabla=True):
    r = np.zeros((9, g.N, g.N, g.nc), order="F")
    n = np.zeros((9, 2, g.N, g.N, g.nc), order="F") if want_nabla else None
    rc = g.L.sd2d_compresid_noalloc(g.h, r.ctypes.data_as(C.c_void_p),
                                    None if n is None else n.ctypes.data_as(C.c_void_p))
    assert rc == abi.SD2D_OK, g.L.sd2d_last_error(g.h).decode()
    return r, n


def _case(nr, nt, N):
    m = M.ogrid(nr, nt, 0.5, 15.0)
    p = M.RunParams(N=N, dt=5e-4)
    case = M.build_case(m, p)
    return case, M.smooth_perturbed_state(case, 0.05, seed=3)


@pytest.mark.parametrize("sizes", [((16, 48, 5), (6, 12, 4))])
def test_noalloc_equals_compresid(sizes):
    from sddc_amd.solver import Sd2d
    for nr, nt, N in sizes:
        print(f"[no-alloc] ogrid {nr}x{nt} N={N}", flush=True)
        case, Q0 = _case(nr, nt, N)
        a, b = Sd2d(case), Sd2d(case)
        try:
            a.set_state(Q0, [0, 0, 0, 0])
            b.set_state(Q0, [0, 0, 0, 0])
            for call in range(2):
                ra, na = a.compresid()
                rb, nb = _noalloc(b)
                assert np.array_equal(ra, rb), (nr, nt, N, call)
                assert np.array_equal(na, nb), (nr, nt, N, call)
                assert np.array_equal(a.maxima(), b.maxima()), (nr, nt, N, call)
                assert a.diag()["c_h"] == b.diag()["c_h"]
            assert b.diag()["c_h"] > 0.0
            # resid alone (no nablaQs copy) is the same too
            rb2, _ = _noalloc(b, want_nabla=False)
            ra2, _ = a.compresid(want_nabla=False)
            assert np.array_equal(ra2, rb2)
        finally:
            a.close()
            b.close()


def test_noalloc_refuses_oversize_mesh():
    from sddc_amd.solver import Sd2d
    print("[no-alloc] oversize mesh", flush=True)
    m = M.ogrid(64, 320, 0.5, 15.0)                    # 20480 cells at N = 5
    case = M.build_case(m, M.RunParams(N=5, dt=5e-4))
    assert 2 * 9 * 25 * case.ncell > abi.SD2D_NOALLOC_MAX_DOUBLES
    g = Sd2d(case)
    try:
        r = np.zeros((9, 5, 5, case.ncell), order="F")
        rc = g.L.sd2d_compresid_noalloc(g.h, r.ctypes.data_as(C.c_void_p), None)
        assert rc == abi.SD2D_EINVAL
        assert "static buffer" in g.L.sd2d_last_error(g.h).decode()
        assert not r.any()                              # nothing ran
    finally:
        g.close()
