"""ORACLE — test infrastructure only.

FP64 CPU restatement of the deterministic atlas log-likelihood for surface subjects
(DeterministicAtlas::UpdateFixedEffectsAndComputeCompleteLogLikelihood,
DeterministicAtlas.cxx:100-151) and of the template-gradient smoothing
(AbstractAtlas::ConvolveGradTemplate, AbstractAtlas.cxx:727-751), composed from the
oracle's Diffeos (deformations.py: Euler shooting, Heun flow) and surface data terms
(surfaces.py), each pinned on its own.  The composition itself is a sum of those pinned
pieces; its gradient is checked against central differences in the GPU tests.
"""
from __future__ import annotations

import numpy as np

from . import deformations as dfo
from . import surfaces as so
from .oracle import ExactKernel


def loglik(cp, momenta, template_V, template_F, targets, *, kind, deformation_width,
           attachment_width, data_sigma_squared, T, scheme="euler", improved_euler=True):
    """(data term, regularity term, residuals) of DeterministicAtlas.cxx:100-151 with the
    RKHS regularity."""
    match = getattr(so, kind + "_match")
    data, reg, res = 0.0, 0.0, []
    for a, (tv, tf) in zip(momenta, targets):
        P, A = dfo.shoot(cp, a, deformation_width, T, scheme=scheme)
        X = dfo.flow(P, A, template_V, deformation_width, T, improved_euler=improved_euler)
        r = match(X[-1], template_F, tv, tf, attachment_width)
        res.append(r)
        data -= r / data_sigma_squared
        k = ExactKernel(cp.shape[1], deformation_width)
        k.SetSources(cp)
        k.SetWeights(a)
        reg -= float(np.sum(k.Convolve(cp) * a))
    return 0.5 * data, 0.5 * reg, np.array(res)


def smooth_template_gradient(template_V, grad_l2, width):
    """Convolve of the L2 template gradient at the smoothing width (AbstractAtlas.cxx:727-751)."""
    if width < 1e-20:
        return grad_l2
    k = ExactKernel(template_V.shape[1], width)
    k.SetSources(template_V)
    k.SetWeights(grad_l2)
    return k.Convolve(template_V)
