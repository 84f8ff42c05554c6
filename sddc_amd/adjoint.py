"""Parameter gradients of the time-averaged wall drag of a cylinder run,
from the device adjoint (sd2d_fwd / sd2d_vjp).

The objective J = (1/nsteps) sum_n D(Q_n) depends on the free-stream Mach
number M and the field angle theta through
  - the boundary ghosts, which read rinf, pinf, uinf, vinf, bxinf, byinf
    (the device returns their cotangents pbar), with uinf = M c_inf,
    |B| = uinf sqrt(rinf) / M_A (or a fixed |B|), bx = |B| cos(theta),
    by = |B| sin(theta);
  - the initial state (the free stream with the wall ramp), whose
    cotangent qbar0 the device returns and whose parameter derivative is
    taken here by central differences of the state builder (exact up to
    rounding: the state is at most quadratic in M and |B|).
The time step is held fixed.  The coefficient of the artificial viscosity
(AV.f90:10) has gradient exactly zero: the sensor does not act on the
state (getrusanov.f90:325-328).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import cases
from . import mesh as M
from .solver import Sd2d


@dataclasses.dataclass
class CylinderSpec:
    nr: int
    ntheta: int
    N: int
    r1: float = 15.0
    mach: float = 1.5
    alfven_mach: float | None = None
    b_angle_deg: float = 0.0
    bmag: float = 0.3
    ramp: float = cases.RAMP
    cfl: float = 0.3


def free_stream(spec: CylinderSpec, mach: float, angle_deg: float):
    """(rinf, pinf, uinf, vinf, bxinf, byinf) of the cylinder runs."""
    gam = M.GAM
    rinf, pinf = 1.0, 1.0 / gam
    u = mach * math.sqrt(gam * pinf / rinf)
    bmag = spec.bmag if spec.alfven_mach is None else u * math.sqrt(rinf) / spec.alfven_mach
    th = math.radians(angle_deg)
    return rinf, pinf, u, 0.0, bmag * math.cos(th), bmag * math.sin(th)


class CylinderObjective:
    """The mesh is built once; J and its gradients for given (M, theta)."""

    def __init__(self, spec: CylinderSpec):
        self.spec = spec
        self.base, Q0 = cases.cylinder(spec.nr, spec.ntheta, spec.N, r1=spec.r1, mach=spec.mach,
                                       alfven_mach=spec.alfven_mach,
                                       b_angle_deg=spec.b_angle_deg, bmag=spec.bmag,
                                       ramp=spec.ramp, cfl=spec.cfl)
        self.dt = self.base.params.dt

    def case(self, mach: float, angle_deg: float):
        r, p, u, v, bx, by = free_stream(self.spec, mach, angle_deg)
        params = dataclasses.replace(self.base.params, rinf=r, pinf=p, uinf=u, vinf=v,
                                     bxinf=bx, byinf=by, dt=self.dt)
        c = dataclasses.replace(self.base, params=params)
        Q0 = M.freestream_state(c, wall_ramp=(0.5, self.spec.ramp))
        return c, Q0

    def value(self, mach: float, angle_deg: float, nsteps: int, ckpt: int) -> float:
        c, Q0 = self.case(mach, angle_deg)
        g = Sd2d(c)
        g.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
        J = g.fwd(nsteps, ckpt)
        g.close()
        return J

    def gradients(self, mach: float, angle_deg: float, nsteps: int, ckpt: int,
                  h_ic: float = 1e-3):
        """J and dJ/dM, dJ/dtheta (theta in radians), dJ/d(AV coefficient)."""
        c, Q0 = self.case(mach, angle_deg)
        g = Sd2d(c)
        g.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
        J = g.fwd(nsteps, ckpt)
        q0b, pb, _ = g.vjp(None, 1.0)
        g.close()
        grads = {}
        for name, dm, da in (("mach", h_ic, 0.0), ("angle_rad", 0.0, math.degrees(h_ic))):
            fp = np.array(free_stream(self.spec, mach + dm, angle_deg + da))
            fm = np.array(free_stream(self.spec, mach - dm, angle_deg - da))
            dp = (fp - fm) / (2 * h_ic)
            Qp = self.case(mach + dm, angle_deg + da)[1]
            Qm = self.case(mach - dm, angle_deg - da)[1]
            dQ = (Qp - Qm) / (2 * h_ic)
            grads[name] = float(np.sum(q0b * dQ) + np.dot(pb, dp))
        grads["av_coefficient"] = 0.0
        return J, grads
