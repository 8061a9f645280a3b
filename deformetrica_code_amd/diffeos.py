"""Host-side mirror of the reference's deformation object for the kernel-sum callers.

``HIPDiffeos`` follows Deformetrica's ``Diffeos`` interface (SetKernelWidth,
SetNumberOfTimePoints, SetT0/SetTN, SetStartPositions/SetStartMomentas,
SetLandmarkPoints, UseImprovedEuler, Update, GetTrajectoryPositions/Momentas,
GetLandmarkPointsTrajectory; src/core/model_tools/deformations/Diffeos.h) with every
kernel sum on the GPU and the trajectories resident in device memory between the
shooting, the flow and the adjoint.  ``GradientOfTerminalCost`` replaces
Diffeos::IntegrateAdjointEquations + AdjointEquationsIntegrator (Diffeos.cxx:1390-1443)
with the exact discrete adjoint of the forward scheme, recomputed from checkpoints.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray, Stream, Workspace


class HIPDiffeos:
    def __init__(self, dim: int, precision: str = "fp64"):
        if dim not in (2, 3):
            raise ValueError("dimension must be 2 or 3")
        self.dim = dim
        self.precision = _lib.precision_code(precision)
        self.kernel_width = 1.0
        self.T = 10           # Diffeos.cxx:26 defaults
        self.t0, self.tN = 0.0, 1.0
        self.improved_euler = True
        self.scheme = _lib.EULER
        self.p0 = None
        self.a0 = None
        self.x0 = None
        self._P = self._A = self._X = None
        self._stream = Stream()
        self._ws = Workspace()

    # --- parameters (Diffeos.h setters)
    def SetKernelWidth(self, h):
        self.kernel_width = float(h)

    def SetNumberOfTimePoints(self, T):
        self.T = int(T)

    def SetT0(self, t0):
        self.t0 = float(t0)

    def SetTN(self, tN):
        self.tN = float(tN)

    def UseImprovedEuler(self):
        self.improved_euler = True

    def UseStandardEuler(self):
        self.improved_euler = False

    def SetShootingScheme(self, scheme: str):
        """"euler" (Diffeos::Shoot) or "heun" (the RK2 variant)."""
        self.scheme = {"euler": _lib.EULER, "heun": _lib.HEUN, "rk2": _lib.HEUN}[scheme]

    def SetStartPositions(self, p0):
        self.p0 = np.ascontiguousarray(p0, dtype=np.float64)

    def SetStartMomentas(self, a0):
        self.a0 = np.ascontiguousarray(a0, dtype=np.float64)

    def SetLandmarkPoints(self, x0):
        self.x0 = np.ascontiguousarray(x0, dtype=np.float64)

    # --- forward
    def Update(self):
        L = _lib.load()
        n, d, T = self.p0.shape[0], self.dim, self.T
        if self.a0.shape != self.p0.shape:
            raise RuntimeError("start positions and momenta differ in shape")
        st = self._stream
        p0 = DeviceArray.from_numpy(self.p0, st)
        a0 = DeviceArray.from_numpy(self.a0, st)
        self._P = DeviceArray((T, n, d))
        self._A = DeviceArray((T, n, d))
        wp, wb = self._ws.get(L.dfm_shoot_workspace(d, n, self.precision))
        _lib.check(L.dfm_shoot(st.handle, d, self.kernel_width, self.precision, self.scheme, p0.ptr,
                               a0.ptr, n, T, self.t0, self.tN, self._P.ptr, self._A.ptr, wp, wb))
        self._X = None
        if self.x0 is not None:
            m = self.x0.shape[0]
            x0 = DeviceArray.from_numpy(self.x0, st)
            self._X = DeviceArray((T, m, d))
            wp, wb = self._ws.get(L.dfm_flow_workspace(d, n, m, self.precision))
            _lib.check(L.dfm_flow(st.handle, d, self.kernel_width, self.precision,
                                  int(self.improved_euler), self._P.ptr, self._A.ptr, n, x0.ptr, m,
                                  T, self.t0, self.tN, self._X.ptr, wp, wb))
        st.synchronize()

    def GetTrajectoryPositions(self):
        return list(self._P.numpy(self._stream))

    def GetTrajectoryMomentas(self):
        return list(self._A.numpy(self._stream))

    def GetLandmarkPointsTrajectory(self):
        return list(self._X.numpy(self._stream))

    def GetDeformedLandmarkPoints(self):
        return self._X.numpy(self._stream)[-1]

    # --- backward
    def GradientOfTerminalCost(self, theta_T, checkpoint_every: int = 5):
        """(dE/dp0, dE/da0, dE/dx0) for dE/dx_final = theta_T."""
        L = _lib.load()
        n, d, T = self.p0.shape[0], self.dim, self.T
        m = self.x0.shape[0]
        st = self._stream
        p0 = DeviceArray.from_numpy(self.p0, st)
        a0 = DeviceArray.from_numpy(self.a0, st)
        x0 = DeviceArray.from_numpy(self.x0, st)
        th = DeviceArray.from_numpy(np.asarray(theta_T, dtype=np.float64), st)
        gp, ga, gx = DeviceArray((n, d)), DeviceArray((n, d)), DeviceArray((m, d))
        wp, wb = self._ws.get(L.dfm_adjoint_workspace(d, n, m, T, int(checkpoint_every),
                                                      self.precision))
        _lib.check(L.dfm_shoot_flow_adjoint(st.handle, d, self.kernel_width, self.precision,
                                            self.scheme, int(self.improved_euler), p0.ptr, a0.ptr, n,
                                            x0.ptr, m, T, self.t0, self.tN, int(checkpoint_every),
                                            th.ptr, gp.ptr, ga.ptr, gx.ptr, wp, wb))
        return gp.numpy(st), ga.numpy(st), gx.numpy(st)


def shoot_vjp(p, a, mu_p, mu_a, sigma, precision="fp64"):
    """Transpose Jacobian of the shooting field at (p, a) applied to (mu_p, mu_a)."""
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = p.shape
    st = Stream()
    dp, da, dmp, dma = (DeviceArray.from_numpy(v, st) for v in (p, a, mu_p, mu_a))
    jp, ja = DeviceArray((n, d)), DeviceArray((n, d))
    ws = Workspace()
    wp, wb = ws.get(L.dfm_workspace_size(_lib.OP_SHOOT_VJP, d, d, n, n, prec))
    _lib.check(L.dfm_shoot_vjp(st.handle, d, float(sigma), prec, dp.ptr, da.ptr, dmp.ptr, dma.ptr,
                               n, jp.ptr, ja.ptr, wp, wb))
    return jp.numpy(st), ja.numpy(st)
