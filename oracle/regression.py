"""ORACLE — test infrastructure only.

FP64 CPU restatement of Deformetrica's geodesic Regression model for landmark objects
(Regression.cxx): one geodesic from the template, observations matched at the time
indices of the time grid, composed from the oracle's Diffeos (deformations.py: shooting,
flow, and the continuous adjoint with jump times) and data terms (surfaces.py, curves.py),
each pinned on its own.

  loglik             UpdateFixedEffectsAndComputeCompleteLogLikelihood (:198-234):
                     [-1/2 sum_j sum_i r_ji / data_sigma_squared_i, -1/2 m^T K(cp, cp) m]
  gradient_reference ComputeCompleteLogLikelihoodGradient (:284-316) through
                     ComputeDataTermGradient (:600-662): each observation's match
                     gradients divided by 2 data_sigma_squared_i, the adjoint with those
                     conditions at the observations' time indices; gradients
                     -xi_pos(0), -xi_mom(0), -theta(0) (smoothed per object).  The
                     regularity gradient is not added: AddGradientRegularityTerm
                     (:664-677) takes its gradients by value, so its updates are lost.

Objects are (kind, cells, attachment_width, data_sigma_squared) as in atlas.py; each
observation is one (points, cells) pair per object.
"""
from __future__ import annotations

import numpy as np

from . import deformations as dfo
from .atlas import _match, _match_gradient, _split, smooth_template_gradient
from .oracle import ExactKernel


def time_index(age, T, t0, tN):
    """The launcher's time index of an observation (estimate_geodesic_regression.cxx:
    229-234): int((age - t0) * ((T - 1) / (tN - t0)) + 0.5f)."""
    return int((age - t0) * ((T - 1) / (tN - t0)) + np.float32(0.5))


def _trajectory(cp, mom, template, deformation_width, T, t0, tN, improved_euler):
    P, A = dfo.shoot(cp, mom, deformation_width, T, t0=t0, tN=tN)
    X = dfo.flow(P, A, np.concatenate(template), deformation_width, T, t0=t0, tN=tN,
                 improved_euler=improved_euler)
    return P, A, X


def loglik(cp, mom, template, objects, observations, time_indices, *, deformation_width, T,
           t0=0.0, tN=1.0, improved_euler=True):
    """(data term, regularity term, residuals (observations x objects))."""
    counts = [len(v) for v in template]
    _, _, X = _trajectory(cp, mom, template, deformation_width, T, t0, tN, improved_euler)
    res = [[_match(kind, cells, x, tv, tf, width)
            for (kind, cells, width, _), x, (tv, tf)
            in zip(objects, _split(X[ti], counts), obs)]
           for obs, ti in zip(observations, time_indices)]
    data = 0.0
    for r in res:
        for (_, _, _, s2), ri in zip(objects, r):
            data -= 0.5 * ri / s2
    k = ExactKernel(cp.shape[1], deformation_width)
    k.SetSources(cp)
    k.SetWeights(mom)
    reg = 0.0
    reg -= float(np.sum(k.Convolve(cp) * mom))
    reg *= 0.5
    return data, reg, np.array(res)


def gradient_reference(cp, mom, template, objects, observations, time_indices, *,
                       deformation_width, T, t0=0.0, tN=1.0, improved_euler=True,
                       smoothing_width=0.0, smoothing_template=None):
    """(d/dcp, d/dm, [d/dT0_i smoothed], [d/dT0_i in L2]) as the reference computes them."""
    counts = [len(v) for v in template]
    P, A, X = _trajectory(cp, mom, template, deformation_width, T, t0, tN, improved_euler)
    conditions = [np.concatenate([
        _match_gradient(kind, cells, x, tv, tf, width) / (2.0 * s2)
        for (kind, cells, width, s2), x, (tv, tf) in zip(objects, _split(X[ti], counts), obs)])
        for obs, ti in zip(observations, time_indices)]
    xp, xm, th0 = dfo.adjoint_reference(P, A, X, conditions, deformation_width, T, t0=t0, tN=tN,
                                        improved_euler=improved_euler,
                                        jump_times=list(time_indices))
    l2 = _split(-th0, counts)
    sources = template if smoothing_template is None else smoothing_template
    sob = [smooth_template_gradient(v, g, smoothing_width) for v, g in zip(sources, l2)]
    return -xp, -xm, sob, l2
