"""ORACLE — test infrastructure only.

FP64 CPU restatement of the deterministic atlas log-likelihood
(DeterministicAtlas::UpdateFixedEffectsAndComputeCompleteLogLikelihood,
DeterministicAtlas.cxx:100-151) and of the template-gradient smoothing
(AbstractAtlas::ConvolveGradTemplate, AbstractAtlas.cxx:727-751) for templates of one or
several landmark objects (DeformableMultiObject: triangle surfaces and polylines, their points
concatenated in object order for the flow), composed from the oracle's Diffeos
(deformations.py: Euler shooting, Euler or Heun flow) and data terms (surfaces.py,
curves.py), each pinned on its own, and of the gradient as the reference assembles it with
its continuous adjoint (``multi_gradient_reference``).  The compositions are sums of those
pinned pieces; the GPU tests check the log-likelihood's gradient against central
differences as well.

An object is (kind, cells, attachment_width, data_sigma_squared): kind "current" or
"varifold" with cells (f x 3) triangles or (e x 2) segments, "pointcloud" with the
template's point weights in place of the cells, or "landmark" (cells and width unused).  A
subject's target is one (points, cells) pair per object, the cells being the target's point
weights for a point cloud.
"""
from __future__ import annotations

import numpy as np

from . import curves as oc
from . import deformations as dfo
from . import points as po
from . import surfaces as so
from .oracle import ExactKernel


def _terms(cells):
    return so if np.asarray(cells).shape[1] == 3 else oc


def _match(kind, cells, x, tv, tf, width):
    if kind == "pointcloud":
        return po.pointcloud_match(x, cells, tv, tf, width)
    if kind == "landmark":
        return po.landmark_match(x, tv)
    return getattr(_terms(cells), kind + "_match")(x, cells, tv, tf, width)


def _match_gradient(kind, cells, x, tv, tf, width):
    if kind == "pointcloud":
        return po.pointcloud_match_gradient(x, cells, tv, tf, width)
    if kind == "landmark":
        return po.landmark_match_gradient(x, tv)
    return getattr(_terms(cells), kind + "_match_gradient")(x, cells, tv, tf, width)


def _split(X, counts):
    return np.split(X, np.cumsum(counts)[:-1])


def multi_loglik(cp, momenta, template, objects, targets, *, deformation_width, T,
                 scheme="euler", improved_euler=True):
    """(data term, regularity term, residuals (subjects x objects)) of
    DeterministicAtlas.cxx:100-151 with the RKHS regularity: the data term is
    -1/2 sum_s sum_i r_si / data_sigma_squared_i."""
    counts = [len(v) for v in template]
    X0 = np.concatenate(template)
    res = []
    for a, target in zip(momenta, targets):
        P, A = dfo.shoot(cp, a, deformation_width, T, scheme=scheme)
        X = dfo.flow(P, A, X0, deformation_width, T, improved_euler=improved_euler)
        res.append([_match(kind, cells, x, tv, tf, width)
                    for (kind, cells, width, _), x, (tv, tf)
                    in zip(objects, _split(X[-1], counts), target)])
    data = 0.0
    for r in res:
        for (_, _, _, s2), ri in zip(objects, r):
            data -= ri / s2
    reg = 0.0
    k = ExactKernel(cp.shape[1], deformation_width)
    k.SetSources(cp)
    for a in momenta:
        k.SetWeights(a)
        reg -= float(np.sum(k.Convolve(cp) * a))
    return 0.5 * data, 0.5 * reg, np.array(res)


def multi_gradient_reference(cp, momenta, template, objects, targets, *, deformation_width, T,
                             smoothing_width=0.0, smoothing_template=None, scheme="euler",
                             improved_euler=True):
    """DeterministicAtlas::ComputeCompleteLogLikelihoodGradient (DeterministicAtlas.cxx:
    159-201, 236-281) with the data term through AbstractAtlas::ComputeDataTermGradientSubject
    (AbstractAtlas.cxx:562-616): per subject each object's match gradient divided by
    2 data_sigma_squared_i, concatenated, the reference's continuous adjoint, data gradients
    summed over subjects, then the RKHS regularity.  The template gradient of each object
    is smoothed with that object's points of ``smoothing_template`` (default: ``template``)
    as sources.  Returns (d/dcp, d/da_s stacked, [d/dT0_i smoothed], [d/dT0_i in L2])."""
    d = cp.shape[1]
    counts = [len(v) for v in template]
    X0 = np.concatenate(template)
    g_cp = np.zeros_like(cp)
    g_mom = np.zeros_like(momenta)
    g_tpl = np.zeros_like(X0)
    for s, (a, target) in enumerate(zip(momenta, targets)):
        P, A = dfo.shoot(cp, a, deformation_width, T, scheme=scheme)
        X = dfo.flow(P, A, X0, deformation_width, T, improved_euler=improved_euler)
        theta = np.concatenate([
            _match_gradient(kind, cells, x, tv, tf, width) / (2.0 * s2)
            for (kind, cells, width, s2), x, (tv, tf) in zip(objects, _split(X[-1], counts), target)])
        xp, xm, th0 = dfo.adjoint_reference(P, A, X, theta, deformation_width, T,
                                            improved_euler=improved_euler)
        g_cp -= xp
        g_mom[s] = -xm
        g_tpl -= th0
    l2 = _split(g_tpl, counts)
    sources = template if smoothing_template is None else smoothing_template
    sob = [smooth_template_gradient(v, g, smoothing_width) for v, g in zip(sources, l2)]
    k = ExactKernel(d, deformation_width)
    k.SetSources(cp)
    for s, a in enumerate(momenta):  # AddGradientRegularityTerm, DeterministicAtlas.cxx:236-262
        k.SetWeights(a)
        g_mom[s] -= k.Convolve(cp)
        g_cp -= k.ConvolveGradient(cp, a)
    return g_cp, g_mom, sob, l2


def loglik(cp, momenta, template_V, template_F, targets, *, kind, deformation_width,
           attachment_width, data_sigma_squared, T, scheme="euler", improved_euler=True):
    """multi_loglik for a one-object template (template_V, template_F) and targets
    [(points, cells)] per subject: (data term, regularity term, residuals per subject)."""
    data, reg, res = multi_loglik(cp, momenta, [template_V],
                                  [(kind, template_F, attachment_width, data_sigma_squared)],
                                  [[t] for t in targets], deformation_width=deformation_width,
                                  T=T, scheme=scheme, improved_euler=improved_euler)
    return data, reg, res[:, 0]


def gradient_reference(cp, momenta, template_V, template_F, targets, *, kind, deformation_width,
                       attachment_width, data_sigma_squared, T, smoothing_width=0.0,
                       scheme="euler", improved_euler=True):
    """multi_gradient_reference for a one-object template: (d/dcp, d/da_s stacked, d/dT0
    smoothed with the template as sources)."""
    g_cp, g_mom, sob, _ = multi_gradient_reference(
        cp, momenta, [template_V], [(kind, template_F, attachment_width, data_sigma_squared)],
        [[t] for t in targets], deformation_width=deformation_width, T=T,
        smoothing_width=smoothing_width, scheme=scheme, improved_euler=improved_euler)
    return g_cp, g_mom, sob[0]


def smooth_template_gradient(template_V, grad_l2, width):
    """Convolve of the L2 template gradient at the smoothing width (AbstractAtlas.cxx:727-751)."""
    if width < 1e-20:
        return grad_l2
    k = ExactKernel(template_V.shape[1], width)
    k.SetSources(template_V)
    k.SetWeights(grad_l2)
    return k.Convolve(template_V)
