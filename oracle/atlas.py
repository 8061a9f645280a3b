"""ORACLE — test infrastructure only.

FP64 CPU restatement of the deterministic atlas log-likelihood for surface subjects
(DeterministicAtlas::UpdateFixedEffectsAndComputeCompleteLogLikelihood,
DeterministicAtlas.cxx:100-151) and of the template-gradient smoothing
(AbstractAtlas::ConvolveGradTemplate, AbstractAtlas.cxx:727-751), composed from the
oracle's Diffeos (deformations.py: Euler shooting, Heun flow) and surface data terms
(surfaces.py), each pinned on its own, and of the gradient as the reference assembles it
with its continuous adjoint (``gradient_reference``).  The compositions are sums of those
pinned pieces; the GPU tests check the log-likelihood's gradient against central
differences as well.
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


def gradient_reference(cp, momenta, template_V, template_F, targets, *, kind, deformation_width,
                       attachment_width, data_sigma_squared, T, smoothing_width=0.0,
                       scheme="euler", improved_euler=True):
    """DeterministicAtlas::ComputeCompleteLogLikelihoodGradient (DeterministicAtlas.cxx:
    159-201, 236-281) with the data term through AbstractAtlas::ComputeDataTermGradientSubject
    (AbstractAtlas.cxx:562-616): the reference's continuous adjoint per subject, data
    gradients summed over subjects, then the RKHS regularity.  Returns (d/dcp, d/da_s
    stacked, d/dT0 smoothed)."""
    grad = getattr(so, kind + "_match_gradient")
    d = cp.shape[1]
    g_cp = np.zeros_like(cp)
    g_mom = np.zeros_like(momenta)
    g_tpl = np.zeros_like(template_V)
    for s, (a, (tv, tf)) in enumerate(zip(momenta, targets)):
        P, A = dfo.shoot(cp, a, deformation_width, T, scheme=scheme)
        X = dfo.flow(P, A, template_V, deformation_width, T, improved_euler=improved_euler)
        theta = grad(X[-1], template_F, tv, tf, attachment_width) / (2.0 * data_sigma_squared)
        xp, xm, th0 = dfo.adjoint_reference(P, A, X, theta, deformation_width, T,
                                            improved_euler=improved_euler)
        g_cp -= xp
        g_mom[s] = -xm
        g_tpl -= th0
    g_tpl = smooth_template_gradient(template_V, g_tpl, smoothing_width)
    k = ExactKernel(d, deformation_width)
    k.SetSources(cp)
    for s, a in enumerate(momenta):  # AddGradientRegularityTerm, DeterministicAtlas.cxx:236-262
        k.SetWeights(a)
        g_mom[s] -= k.Convolve(cp)
        g_cp -= k.ConvolveGradient(cp, a)
    return g_cp, g_mom, g_tpl


def smooth_template_gradient(template_V, grad_l2, width):
    """Convolve of the L2 template gradient at the smoothing width (AbstractAtlas.cxx:727-751)."""
    if width < 1e-20:
        return grad_l2
    k = ExactKernel(template_V.shape[1], width)
    k.SetSources(template_V)
    k.SetWeights(grad_l2)
    return k.Convolve(template_V)
