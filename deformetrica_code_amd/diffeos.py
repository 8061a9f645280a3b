"""Host-side mirror of the reference's deformation object for the kernel-sum callers.

``HIPDiffeos`` follows Deformetrica's ``Diffeos`` interface (SetKernelWidth,
SetNumberOfTimePoints, SetT0/SetTN, SetStartPositions/SetStartMomentas,
SetLandmarkPoints, UseImprovedEuler, Update, GetTrajectoryPositions/Momentas,
GetLandmarkPointsTrajectory; src/core/model_tools/deformations/Diffeos.h) with every
kernel sum on the GPU and the trajectories resident in device memory between the
shooting, the flow and the adjoint.  ``IntegrateAdjointEquations`` is the reference's
Diffeos::IntegrateAdjointEquations + AdjointEquationsIntegrator (Diffeos.cxx:1390-1443):
the continuous adjoint equations integrated backward; ``GradientOfTerminalCost`` is the
exact discrete adjoint of the forward scheme (equal to O(dt)); both recompute the forward
states from checkpoints.
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
        self._bufs = {}

    def _buf(self, name, shape):
        """A device array kept between calls (reallocated only when the shape changes), so
        repeated Update calls allocate nothing."""
        b = self._bufs.get(name)
        if b is None or b.shape != tuple(shape):
            b = self._bufs[name] = DeviceArray(shape)
        return b

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
        # the start positions and momenta in one buffer, one copy queued ahead of the shoot
        # (the host array lives until the synchronisation below)
        pa0 = self._buf("pa0", (2, n, d))
        host_pa0 = np.ascontiguousarray(np.stack([self.p0, self.a0]), dtype=np.float64)
        pa0.upload_async(host_pa0, st)
        try:
            p0_ptr, a0_ptr = pa0.ptr, pa0.ptr + n * d * 8
            self._P = self._buf("P", (T, n, d))
            self._A = self._buf("A", (T, n, d))
            wp, wb = self._ws.get(L.dfm_shoot_workspace(d, n, self.precision))
            _lib.check(L.dfm_shoot(st.handle, d, self.kernel_width, self.precision, self.scheme,
                                   p0_ptr, a0_ptr, n, T, self.t0, self.tN, self._P.ptr,
                                   self._A.ptr, wp, wb))
            self._X = None
            if self.x0 is not None:
                m = self.x0.shape[0]
                x0 = self._buf("x0", self.x0.shape)
                x0.upload(self.x0, st)
                self._X = self._buf("X", (T, m, d))
                wp, wb = self._ws.get(L.dfm_flow_workspace(d, n, m, self.precision))
                _lib.check(L.dfm_flow(st.handle, d, self.kernel_width, self.precision,
                                      int(self.improved_euler), self._P.ptr, self._A.ptr, n,
                                      x0.ptr, m, T, self.t0, self.tN, self._X.ptr, wp, wb))
        finally:
            st.synchronize()  # also on an error: the queued copy still reads host_pa0

    def GetTrajectoryPositions(self):
        return list(self._P.numpy(self._stream))

    def GetTrajectoryMomentas(self):
        return list(self._A.numpy(self._stream))

    def GetLandmarkPointsTrajectory(self):
        return list(self._X.numpy(self._stream))

    def GetDeformedLandmarkPoints(self):
        return self._X.numpy(self._stream)[-1]

    # --- parallel transport
    def ParallelTransport(self, initial_momenta, initial_cp, initial_time, target_times):
        """Diffeos::ParallelTransport (Diffeos.cxx:293-582) along the shot geodesic: the
        transported momenta (list, one n x d array per target time)."""
        L = _lib.load()
        n, d, T = self.p0.shape[0], self.dim, self.T
        st = self._stream
        mom = DeviceArray.from_numpy(initial_momenta, st)
        cp = DeviceArray.from_numpy(initial_cp, st)
        tt = np.ascontiguousarray(target_times, dtype=np.float64)
        out = DeviceArray((len(tt), n, d))
        wp, wb = self._ws.get(L.dfm_parallel_transport_workspace(d, n, cp.shape[0], self.precision))
        _lib.check(L.dfm_parallel_transport(
            st.handle, d, self.kernel_width, self.precision, self._P.ptr, self._A.ptr, n, T,
            self.t0, self.tN, mom.ptr, cp.ptr, cp.shape[0], float(initial_time),
            tt.ctypes.data_as(ctypes.c_void_p), len(tt), out.ptr, wp, wb))
        return list(out.numpy(st))

    # --- backward
    def GradientOfTerminalCost(self, theta_T, checkpoint_every: int = 0):
        """(dE/dp0, dE/da0, dE/dx0) for dE/dx_final = theta_T: the exact discrete adjoint."""
        return self._adjoint(theta_T, checkpoint_every, "dfm_shoot_flow_adjoint")

    def IntegrateAdjointEquations(self, theta_T, checkpoint_every: int = 0, jump_times=None):
        """Diffeos::IntegrateAdjointEquations (Diffeos.cxx:1374-1443) for the landmark
        points: (xi_pos(0), xi_mom(0), theta(0)) of the reference's continuous adjoint
        equations for the terminal landmark cotangent theta_T; the reference's
        GetAdjointPosAt0 / GetAdjointMomAt0 are the first two.  With ``jump_times`` (the
        geodesic regression's time indices), theta_T is the list of landmark conditions at
        those indices, applied as the reference's AdjointEquationsIntegrator applies them
        (dfm_shoot_flow_adjoint_jumps)."""
        if jump_times is None:
            return self._adjoint(theta_T, checkpoint_every, "dfm_shoot_flow_adjoint_equations")
        times = np.ascontiguousarray(jump_times, dtype=np.int32)
        conditions = np.ascontiguousarray(np.stack([np.asarray(c, dtype=np.float64)
                                                    for c in theta_T]))
        if conditions.shape[0] != times.shape[0]:
            raise ValueError("one condition per jump time")
        return self._adjoint(conditions, checkpoint_every, "dfm_shoot_flow_adjoint_jumps",
                             times)

    def _adjoint(self, theta_T, checkpoint_every, entry, jump_times=None):
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
        args = [st.handle, d, self.kernel_width, self.precision, self.scheme,
                int(self.improved_euler), p0.ptr, a0.ptr, n, x0.ptr, m, T, self.t0, self.tN,
                int(checkpoint_every)]
        if jump_times is not None:
            args += [len(jump_times), jump_times.ctypes.data_as(ctypes.c_void_p)]
        _lib.check(getattr(L, entry)(*args, th.ptr, gp.ptr, ga.ptr, gx.ptr, wp, wb))
        return gp.numpy(st), ga.numpy(st), gx.numpy(st)


def convolve_inverse(y, x, sigma, stream=None):
    """K(y, y)^-1 x in FP64 (ExactKernel::ConvolveInverse), on host arrays."""
    L = _lib.load()
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    n, d = y.shape
    dv = x.shape[1]
    st = stream or Stream()
    dy, dx, dv_out = DeviceArray.from_numpy(y, st), DeviceArray.from_numpy(x, st), DeviceArray(x.shape)
    ws = Workspace()
    wp, wb = ws.get(L.dfm_convolve_inverse_workspace(d, dv, n))
    _lib.check(L.dfm_convolve_inverse(st.handle, d, dv, float(sigma), dy.ptr, n, dx.ptr, dv_out.ptr,
                                      wp, wb))
    return dv_out.numpy(st)


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


def _offsets(counts):
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    off[1:] = np.cumsum(counts)
    return off


class HIPBatchDiffeos(HIPDiffeos):
    """Many subjects (each its own control points, momenta and landmarks, with uneven
    counts) shot, flowed and back-propagated together: every kernel sum of a time step is
    one launch over all subjects (dfm_batch_shoot / dfm_batch_flow /
    dfm_batch_shoot_flow_adjoint).  Per-subject setters take lists of arrays; getters
    return lists.  Replaces the per-subject Diffeos clones of the reference's atlas loop
    (AbstractAtlas.cxx:490-560)."""

    def SetStartPositions(self, p0_list):
        self.p0_list = [np.ascontiguousarray(p, dtype=np.float64) for p in p0_list]

    def SetStartMomentas(self, a0_list):
        self.a0_list = [np.ascontiguousarray(a, dtype=np.float64) for a in a0_list]

    def SetLandmarkPoints(self, x0_list):
        self.x0_list = [np.ascontiguousarray(x, dtype=np.float64) for x in x0_list]

    def _segments(self):
        d = self.dim
        if len(self.p0_list) != len(self.a0_list) or any(
                p.shape != a.shape for p, a in zip(self.p0_list, self.a0_list)):
            raise RuntimeError("start positions and momenta differ in shape")
        self.cp_off = _offsets([p.shape[0] for p in self.p0_list])
        self.lm_off = _offsets([x.shape[0] for x in self.x0_list])
        cat = lambda lst: (np.concatenate(lst).reshape(-1, d) if lst else np.zeros((0, d)))
        return cat(self.p0_list), cat(self.a0_list), cat(self.x0_list)

    @staticmethod
    def _cp(off):
        return off.ctypes.data_as(ctypes.c_void_p)

    def Update(self):
        L = _lib.load()
        d, T, S = self.dim, self.T, len(self.p0_list)
        p0, a0, x0 = self._segments()
        st = self._stream
        dp0, da0, dx0 = (DeviceArray.from_numpy(v, st) for v in (p0, a0, x0))
        n, m = p0.shape[0], x0.shape[0]
        self._P, self._A, self._X = DeviceArray((T, n, d)), DeviceArray((T, n, d)), DeviceArray((T, m, d))
        cp, lm = self._cp(self.cp_off), self._cp(self.lm_off)
        wp, wb = self._ws.get(L.dfm_batch_shoot_workspace(d, S, cp, self.precision))
        _lib.check(L.dfm_batch_shoot(st.handle, d, self.kernel_width, self.precision, self.scheme, S,
                                     cp, dp0.ptr, da0.ptr, T, self.t0, self.tN, self._P.ptr,
                                     self._A.ptr, wp, wb))
        wp, wb = self._ws.get(L.dfm_batch_flow_workspace(d, S, cp, lm, self.precision))
        _lib.check(L.dfm_batch_flow(st.handle, d, self.kernel_width, self.precision,
                                    int(self.improved_euler), S, cp, lm, self._P.ptr, self._A.ptr,
                                    dx0.ptr, T, self.t0, self.tN, self._X.ptr, wp, wb))
        st.synchronize()

    def _split(self, arr, off):
        return [arr[..., off[s]:off[s + 1], :] for s in range(len(off) - 1)]

    def GetTrajectoryPositions(self):
        return self._split(self._P.numpy(self._stream), self.cp_off)

    def GetTrajectoryMomentas(self):
        return self._split(self._A.numpy(self._stream), self.cp_off)

    def GetLandmarkPointsTrajectory(self):
        return self._split(self._X.numpy(self._stream), self.lm_off)

    def GetDeformedLandmarkPoints(self):
        return self._split(self._X.numpy(self._stream)[-1], self.lm_off)

    def GradientOfTerminalCost(self, theta_list, checkpoint_every: int = 0):
        """Per-subject lists (dE/dp0, dE/da0, dE/dx0) for dE/dx_final = theta_list[s]."""
        L = _lib.load()
        d, T, S = self.dim, self.T, len(self.p0_list)
        p0, a0, x0 = self._segments()
        th = np.concatenate([np.asarray(t, dtype=np.float64) for t in theta_list]).reshape(-1, d)
        st = self._stream
        dp0, da0, dx0, dth = (DeviceArray.from_numpy(v, st) for v in (p0, a0, x0, th))
        gp, ga, gx = DeviceArray(p0.shape), DeviceArray(p0.shape), DeviceArray(x0.shape)
        cp, lm = self._cp(self.cp_off), self._cp(self.lm_off)
        wp, wb = self._ws.get(L.dfm_batch_adjoint_workspace(d, S, cp, lm, T, int(checkpoint_every),
                                                            self.precision))
        _lib.check(L.dfm_batch_shoot_flow_adjoint(
            st.handle, d, self.kernel_width, self.precision, self.scheme, int(self.improved_euler),
            S, cp, lm, dp0.ptr, da0.ptr, dx0.ptr, T, self.t0, self.tN, int(checkpoint_every),
            dth.ptr, gp.ptr, ga.ptr, gx.ptr, wp, wb))
        return (self._split(gp.numpy(st), self.cp_off), self._split(ga.numpy(st), self.cp_off),
                self._split(gx.numpy(st), self.lm_off))
