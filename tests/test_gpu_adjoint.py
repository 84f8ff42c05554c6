"""Adjoint of the residual and of RK_S5P3 steps (sd2d_compresid_vjp,
sd2d_fwd / sd2d_vjp) against central finite differences of the device's
own forward path, through dot products with seeded directions.

Residual: <Rbar, (R(Q+hV) - R(Q-hV))/2h> = <Qbar, V> and likewise for c_h
and the six free-stream values the boundary ghosts read, to 1e-6 relative
(h = 1e-5 relative to each variable's scale: the FD truncation error is
~1e-10, its rounding ~1e-11).
"""
import numpy as np
import pytest

from tests.util import ORDERED, make_case

pytestmark = pytest.mark.gpu

FIELDS = ["rinf", "pinf", "uinf", "vinf", "bxinf", "byinf"]


def _gpu(case):
    from sddc_amd.solver import Sd2d
    return Sd2d(case)


@pytest.mark.parametrize("name", ORDERED)
def test_residual_vjp_against_fd(name):
    print(f"[gpu-adj] residual {name}", flush=True)
    import copy
    case, Q0 = make_case(name)
    eg, ug = 2.3, 0.7
    mx = [eg, ug, eg, ug]
    rng = np.random.default_rng(5)
    rbar = np.asfortranarray(rng.standard_normal(Q0.shape))
    scale = np.abs(Q0).max(axis=(1, 2, 3))[:, None, None, None]
    V = np.asfortranarray(rng.standard_normal(Q0.shape)) * scale
    g = _gpu(case)

    def R(Q, m=mx, gg=None):
        gg = gg or g
        gg.set_state(Q, m)
        return gg.compresid(want_nabla=False)[0]

    g.set_state(Q0, mx)
    qb, chb, pb = g.compresid_vjp(rbar)
    h = 1e-5
    fd = np.sum(rbar * (R(Q0 + h * V) - R(Q0 - h * V))) / (2 * h)
    ad = np.sum(qb * V)
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    # c_h = sqrt(eg (eg - ug)) through eigvmax_g
    ch = np.sqrt(eg * (eg - ug))
    he = 1e-6
    fdc = np.sum(rbar * (R(Q0, [eg, ug, eg + he, ug]) - R(Q0, [eg, ug, eg - he, ug]))) / (2 * he)
    adc = chb * (2 * eg - ug) / (2 * ch)
    assert abs(adc - fdc) <= 1e-6 * abs(fdc), (adc, fdc)
    for j, f in enumerate(FIELDS):
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        hp = 1e-6
        setattr(cp.params, f, getattr(cp.params, f) + hp)
        setattr(cm.params, f, getattr(cm.params, f) - hp)
        fdp = np.sum(rbar * (R(Q0, gg=_gpu(cp)) - R(Q0, gg=_gpu(cm)))) / (2 * hp)
        assert abs(pb[j] - fdp) <= 1e-6 * max(abs(fdp), 1e-8 * np.abs(rbar).sum()), (f, pb[j], fdp)
    g.close()


@pytest.mark.parametrize("name", ["ogrid4x8_N4", "channel5x4_N3", "periodic4x4_N5"])
def test_step_vjp_against_fd(name):
    """Three RK_S5P3 steps with checkpoints every two steps: the VJP of a
    seeded N(0,1) cotangent on the final state plus the wall-drag
    objective, against central differences of the device forward."""
    print(f"[gpu-adj] steps {name}", flush=True)
    case, Q0 = make_case(name)
    rng = np.random.default_rng(7)
    C = np.asfortranarray(rng.standard_normal(Q0.shape))
    scale = np.abs(Q0).max(axis=(1, 2, 3))[:, None, None, None]
    V = np.asfortranarray(rng.standard_normal(Q0.shape)) * scale
    nsteps, jbar = 3, 2.0
    g = _gpu(case)

    def F(Q):
        g.set_state(Q, [0, 0, 0, 0])
        J = g.fwd(nsteps, 2)
        return np.sum(C * g.get_state()[0]) + jbar * J

    g.set_state(Q0, [0, 0, 0, 0])
    g.fwd(nsteps, 2)
    qb0, pb, ch = g.vjp(C, jbar, want_chbar=True)
    h = 1e-6
    fd = (F(Q0 + h * V) - F(Q0 - h * V)) / (2 * h)
    ad = np.sum(qb0 * V)
    print(f"[gpu-adj] {name}: adjoint {ad:.12e} fd {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}",
          flush=True)
    assert abs(ad - fd) <= 1e-5 * abs(fd), (ad, fd)
    g.close()


@pytest.mark.parametrize("name", ["ogrid4x8_N4", "ogrid_curved5x12_N5"])
def test_adjoint_checkpoints_restart(tmp_path, name):
    """The forward's checkpoints saved to files (sd2d_adj_save) and loaded
    into a fresh context (sd2d_adj_load) give the backward pass of the
    context that ran the forward bit for bit; every checkpoint file is a
    reference restart file (the first one holds the initial state), and a
    context of another mesh refuses the files."""
    print(f"[gpu-adj] checkpoint files {name}", flush=True)
    from sddc_amd import mesh as M
    from sddc_amd.solver import Sd2dError
    case, Q0 = make_case(name)
    C = np.asfortranarray(np.random.default_rng(11).standard_normal(Q0.shape))
    nsteps, ckpt = 7, 3                      # checkpoints at steps 0, 3, 6 and the final state
    g = _gpu(case)
    g.set_state(Q0, [0, 0, 0, 0])
    g.set_time(0.25, 40)
    J = g.fwd(nsteps, ckpt)
    pre = str(tmp_path / "adj")
    g.adj_save(pre)
    qa, pa, ca = g.vjp(C, 2.0, want_chbar=True)
    Qa, ma = g.get_state()
    da = g.diag()
    g.close()
    q0, t0, it0 = M.read_restart(pre + ".0000.dat", case.N)
    assert np.array_equal(q0, Q0) and t0 == 0.25 and it0 == 40
    for k in range(4):
        assert (tmp_path / f"adj.{k:04d}.dat.max").exists()
    h = _gpu(case)
    h.adj_load(pre)
    Qh, mh = h.get_state()
    assert np.array_equal(Qh, Qa) and np.array_equal(mh, ma)
    assert h.diag()["iter"] == da["iter"] == 47 and h.diag()["ctime"] == da["ctime"]
    qb, pbb, cb = h.vjp(C, 2.0, want_chbar=True)
    assert np.array_equal(qa, qb) and np.array_equal(pa, pbb) and np.array_equal(ca, cb)
    h.close()
    other, _ = make_case("channel5x4_N3")
    o = _gpu(other)
    with pytest.raises(Sd2dError, match="another mesh"):
        o.adj_load(pre)
    o.close()
    assert np.isfinite(J)


def test_parameter_gradients_against_fd():
    """dJ/dMach and dJ/dtheta of the time-averaged wall drag (boundary
    ghosts and initial state) against central differences with h = 1e-6;
    the AV coefficient's gradient is exactly 0 (the sensor does not act)."""
    print("[gpu-adj] parameter gradients", flush=True)
    from sddc_amd.adjoint import CylinderObjective, CylinderSpec
    o = CylinderObjective(CylinderSpec(16, 32, 4))
    mach, ang, nsteps, ckpt = 1.5, 10.0, 20, 5
    J, gr = o.gradients(mach, ang, nsteps, ckpt)
    h = 1e-6                          # Mach, and the angle in radians
    hd = np.degrees(h)
    fdm = (o.value(mach + h, ang, nsteps, ckpt) - o.value(mach - h, ang, nsteps, ckpt)) / (2 * h)
    fda = (o.value(mach, ang + hd, nsteps, ckpt) - o.value(mach, ang - hd, nsteps, ckpt)) / (2 * h)
    print(f"[gpu-adj] J={J:.10e} dJ/dM adj {gr['mach']:.10e} fd {fdm:.10e}; "
          f"dJ/dtheta adj {gr['angle_rad']:.10e} fd {fda:.10e}", flush=True)
    assert abs(gr["mach"] - fdm) <= 1e-5 * abs(fdm)
    assert abs(gr["angle_rad"] - fda) <= 1e-5 * abs(fda)
    assert gr["av_coefficient"] == 0.0


def test_torch_autograd_wrapper():
    """The verification harness's autograd Function over sd2d_fwd/sd2d_vjp:
    q0.grad from J.backward() against a central difference of J."""
    print("[gpu-adj] torch wrapper (importing torch)", flush=True)
    import torch
    from tests.adjoint_torch import DeviceDragObjective
    case, Q0 = make_case("ogrid4x8_N4")
    g = _gpu(case)
    q0 = torch.tensor(np.ascontiguousarray(Q0), dtype=torch.float64, requires_grad=True)
    J = DeviceDragObjective.apply(q0, g, 6, 3)
    J.backward()
    rng = np.random.default_rng(11)
    V = rng.standard_normal(Q0.shape) * np.abs(Q0).max(axis=(1, 2, 3))[:, None, None, None]
    h = 1e-6
    Jp = DeviceDragObjective.apply(torch.tensor(np.ascontiguousarray(Q0 + h * V)), g, 6, 3).item()
    Jm = DeviceDragObjective.apply(torch.tensor(np.ascontiguousarray(Q0 - h * V)), g, 6, 3).item()
    fd = (Jp - Jm) / (2 * h)
    ad = float(np.sum(q0.grad.numpy() * V))
    print(f"[gpu-adj] torch: adjoint {ad:.10e} fd {fd:.10e}", flush=True)
    assert abs(ad - fd) <= 1e-5 * abs(fd)
    g.close()
