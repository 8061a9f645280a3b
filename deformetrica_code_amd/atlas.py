"""Deterministic atlas of surfaces or polylines on the GPU: the log-likelihood and gradient of
Deformetrica's DeterministicAtlas (DeterministicAtlas.cxx:100-281, AbstractAtlas.cxx:490-751)
for many subjects at once, through libdfm_hip.so's dfm_atlas_* entry points.

``HIPDeterministicAtlas`` is created once per (template topology, targets, parameters);
``ComputeLogLikelihood`` / ``ComputeCompleteLogLikelihoodGradient`` then run on
device-resident control points, momenta and template vertices.  ``capture()`` records one
gradient evaluation in a HIP graph for replay.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray, Stream


class AtlasDesc(ctypes.Structure):
    """Mirror of dfm_atlas_desc (include/dfm_hip.h)."""
    _fields_ = [
        ("kind", ctypes.c_int), ("precision", ctypes.c_int), ("scheme", ctypes.c_int),
        ("improved_euler", ctypes.c_int), ("number_of_time_points", ctypes.c_int),
        ("checkpoint_every", ctypes.c_int), ("t0", ctypes.c_double), ("tN", ctypes.c_double),
        ("deformation_width", ctypes.c_double), ("attachment_width", ctypes.c_double),
        ("data_sigma_squared", ctypes.c_double), ("smoothing_width", ctypes.c_double),
        ("n_control_points", ctypes.c_long), ("n_template_vertices", ctypes.c_long),
        ("n_template_faces", ctypes.c_long), ("template_faces", ctypes.c_void_p),
        ("nsubjects", ctypes.c_int), ("target_vertex_offsets", ctypes.c_void_p),
        ("target_face_offsets", ctypes.c_void_p), ("target_vertices", ctypes.c_void_p),
        ("target_faces", ctypes.c_void_p), ("adjoint", ctypes.c_int),
        ("dimension", ctypes.c_int), ("cell_corners", ctypes.c_int),
    ]


def _offsets(counts):
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    off[1:] = np.cumsum(counts)
    return off


class HIPDeterministicAtlas:
    """targets: list of (vertices (v_s x d), cells (f_s x c)) per subject, with the template's
    cells (triangles, c = 3, d = 3; or polyline segments, c = 2, d = 2 or 3); d is taken
    from the targets' vertices."""

    def __init__(self, template_faces, n_template_vertices, n_control_points, targets, *,
                 kind="varifold", precision="fp64", deformation_width=1.0,
                 attachment_width=1.0, data_sigma_squared=1.0, smoothing_width=0.0,
                 number_of_time_points=11, t0=0.0, tN=1.0, scheme="euler", improved_euler=True,
                 checkpoint_every=0, adjoint="discrete"):
        L = _lib.load()
        tf = np.ascontiguousarray(template_faces, dtype=np.int32)
        corners = tf.shape[1]
        dim = np.asarray(targets[0][0]).shape[1]
        self.dim = dim
        tv = np.ascontiguousarray(np.concatenate([np.asarray(v, np.float64).reshape(-1, dim)
                                                  for v, _ in targets]))
        tfa = np.ascontiguousarray(np.concatenate([np.asarray(f, np.int32).reshape(-1, corners)
                                                   for _, f in targets]))
        vo = _offsets([len(v) for v, _ in targets])
        fo = _offsets([len(f) for _, f in targets])
        self._keep = (tf, tv, tfa, vo, fo)
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        d = AtlasDesc()
        d.kind = {"current": _lib.CURRENT, "varifold": _lib.VARIFOLD}[kind]
        d.precision = _lib.precision_code(precision)
        d.scheme = {"euler": _lib.EULER, "heun": _lib.HEUN, "rk2": _lib.HEUN}[scheme]
        d.improved_euler = int(improved_euler)
        d.number_of_time_points = int(number_of_time_points)
        d.checkpoint_every = int(checkpoint_every)
        d.t0, d.tN = float(t0), float(tN)
        d.deformation_width = float(deformation_width)
        d.attachment_width = float(attachment_width)
        d.data_sigma_squared = float(data_sigma_squared)
        d.smoothing_width = float(smoothing_width)
        d.n_control_points = int(n_control_points)
        d.n_template_vertices = int(n_template_vertices)
        d.n_template_faces = tf.shape[0]
        d.template_faces = vp(tf)
        d.nsubjects = len(targets)
        d.target_vertex_offsets, d.target_face_offsets = vp(vo), vp(fo)
        d.target_vertices, d.target_faces = vp(tv), vp(tfa)
        d.adjoint = _lib.adjoint_code(adjoint)
        d.dimension, d.cell_corners = int(dim), int(corners)
        self.desc = d
        self.S, self.n, self.m = len(targets), int(n_control_points), int(n_template_vertices)
        self._h = ctypes.c_void_p()
        _lib.check(L.dfm_atlas_create(ctypes.byref(d), ctypes.byref(self._h)))
        self.stream = Stream()
        self.loglik = DeviceArray((2,))
        self.residuals = DeviceArray((self.S,))
        self.grad_cp = DeviceArray((self.n, dim))
        self.grad_momenta = DeviceArray((self.S, self.n, dim))
        self.grad_template = DeviceArray((self.m, dim))

    def device_bytes(self):
        return int(_lib.load().dfm_atlas_device_bytes(self._h))

    def gradient_device(self, cp, momenta, template, with_gradient=True):
        """One evaluation on DeviceArrays, stream-ordered (no synchronisation)."""
        L = _lib.load()
        g = (lambda a: a.ptr) if with_gradient else (lambda a: None)
        _lib.check(L.dfm_atlas_gradient(self._h, self.stream.handle, cp.ptr, momenta.ptr,
                                        template.ptr, self.loglik.ptr, self.residuals.ptr,
                                        g(self.grad_cp), g(self.grad_momenta),
                                        g(self.grad_template)))

    def _upload(self, cp, momenta, template):
        st = self.stream
        return (DeviceArray.from_numpy(cp, st), DeviceArray.from_numpy(momenta, st),
                DeviceArray.from_numpy(template, st))

    def ComputeLogLikelihood(self, cp, momenta, template):
        """(data term, regularity term) and the per-subject residuals."""
        args = self._upload(cp, momenta, template)
        self.gradient_device(*args, with_gradient=False)
        self.stream.synchronize()
        return self.loglik.numpy(self.stream), self.residuals.numpy(self.stream)

    def ComputeCompleteLogLikelihoodGradient(self, cp, momenta, template):
        """loglik terms, d/d cp, d/d momenta (S x n x 3), d/d template vertices."""
        args = self._upload(cp, momenta, template)
        self.gradient_device(*args)
        self.stream.synchronize()
        st = self.stream
        return (self.loglik.numpy(st), self.grad_cp.numpy(st), self.grad_momenta.numpy(st),
                self.grad_template.numpy(st))

    def capture(self, cp, momenta, template):
        """A HIP graph of one gradient evaluation on the given DeviceArrays."""
        L = _lib.load()
        _lib.check(L.dfm_graph_begin_capture(self.stream.handle))
        try:
            self.gradient_device(cp, momenta, template)
        finally:
            g = ctypes.c_void_p()
            rc = L.dfm_graph_end_capture(self.stream.handle, ctypes.byref(g))
        _lib.check(rc)
        return Graph(g, self.stream)

    def __del__(self):
        try:
            if self._h.value:
                _lib.load().dfm_atlas_destroy(self._h)
        except Exception:
            pass


class Graph:
    def __init__(self, handle, stream):
        self._g, self.stream = handle, stream

    def launch(self):
        _lib.check(_lib.load().dfm_graph_launch(self._g, self.stream.handle))

    def __del__(self):
        try:
            if self._g.value:
                _lib.load().dfm_graph_destroy(self._g)
        except Exception:
            pass


def initialize_control_points(bounding_box, spacing):
    """AbstractAtlas::InitializeControlPoints without the image criterion
    (AbstractAtlas.cxx:240-362): a lattice of the given spacing centred in the template's
    bounding box ((d x 2): min, max per axis), x outermost, coordinates accumulated by
    repeated addition as the reference's loops do."""
    box = np.asarray(bounding_box, dtype=np.float64)
    lo, hi = box[:, 0], box[:, 1]
    sp = float(spacing)
    off = 0.5 * (hi - lo - sp * np.floor((hi - lo) / sp))
    axes = []
    for k in range(box.shape[0]):
        vals, v = [], lo[k] + off[k]
        while v <= hi[k]:
            vals.append(v)
            v += sp
        axes.append(vals)
    grid = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.ravel() for g in grid], axis=1)


def surface_bounding_box(vertices, faces):
    """Bounding box of a triangle mesh as OrientedSurfaceMesh / NonOrientedSurfaceMesh keep
    it: over the triangle centers (OrientedSurfaceMesh.cxx:113-115, UpdateBoundingBox)."""
    v = np.asarray(vertices, dtype=np.float64)
    f = np.asarray(faces)
    c = (v[f[:, 0]] + v[f[:, 1]] + v[f[:, 2]]) / 3.0
    return np.stack([c.min(axis=0), c.max(axis=0)], axis=1)


def curve_bounding_box(points, segments):
    """Bounding box of a polyline as OrientedPolyLine / NonOrientedPolyLine keep it: over the
    segment centres (OrientedPolyLine.cxx:299-321, NonOrientedPolyLine.cxx:327-351)."""
    v = np.asarray(points, dtype=np.float64)
    e = np.asarray(segments)
    c = (v[e[:, 0]] + v[e[:, 1]]) / 2.0
    return np.stack([c.min(axis=0), c.max(axis=0)], axis=1)


def point_bounding_box(points):
    """Bounding box of a Landmark or PointCloud object: over its points
    (Landmark::UpdateBoundingBox, Landmark.cxx:371-387)."""
    v = np.asarray(points, dtype=np.float64)
    return np.stack([v.min(axis=0), v.max(axis=0)], axis=1)


def union_bounding_box(boxes):
    """Bounding box of a multi-object template: the union of its objects' boxes
    (DeformableMultiObject::UpdateBoundingBox, DeformableMultiObject.cxx:410-429)."""
    boxes = [np.asarray(b, dtype=np.float64) for b in boxes]
    out = boxes[0].copy()
    for b in boxes[1:]:
        out[:, 0] = np.minimum(out[:, 0], b[:, 0])
        out[:, 1] = np.maximum(out[:, 1], b[:, 1])
    return out


def _copy_effects(effects):
    return {k: ([np.array(x, dtype=np.float64) for x in v] if isinstance(v, (list, tuple))
                else np.array(v, dtype=np.float64)) for k, v in effects.items()}


class HIPAtlasModel:
    """DeterministicAtlas as the estimators see it (AbstractStatisticalModel's
    Get/SetFixedEffects, UpdateFixedEffectsAndComputeCompleteLogLikelihood,
    ComputeCompleteLogLikelihoodGradient; DeterministicAtlas.cxx:100-201) over a
    HIPDeterministicAtlas created with smoothing_width 0.  Fixed effects "ControlPoints"
    (n x d) and "TemplateData" ([vertices]); individual random effects "Momenta" (one
    n x d per subject), d the atlas's dimension.

    The reference's behaviour is kept where it shapes the numbers:
    - SetFixedEffects (AbstractAtlas.cxx:52-63) stores the new fixed effects but moves the
      template geometry to the template data it held *before* the call, so the geometry the
      shot and the data term use lags one call behind the "TemplateData" fixed effect
      (frozen template: the fixed effect is restored and the geometry stays);
    - the template gradient is smoothed with the "TemplateData" fixed effect as sources and
      targets (AbstractAtlas::ConvolveGradTemplate, AbstractAtlas.cxx:727-751), here by one
      GPU Convolve at ``smoothing_width`` (<= 1e-20: none);
    - gradient blocks "ControlPoints" unless frozen, "TemplateData" (smoothed) unless
      frozen, "Momenta"; squared norms in that order, the template's as <L2, smoothed>
      (DeterministicAtlas.cxx:197-200).
    A registration is the one-subject case with a frozen template (estimate_atlas.cxx:
    229-233).  The bounding-box test of the reference's Diffeos is not made: out_of_box is
    always False."""

    def __init__(self, atlas, fixed_effects, *, freeze_template=False,
                 freeze_control_points=False, smoothing_width=0.0):
        if atlas.desc.smoothing_width > 1e-20:
            raise ValueError("create the atlas with smoothing_width 0: the model smooths the "
                             "template gradient itself")
        self.atlas = atlas
        self.freeze_template = bool(freeze_template)
        self.freeze_control_points = bool(freeze_control_points)
        self.smoothing_width = float(smoothing_width)
        self._fixed = _copy_effects(fixed_effects)
        self._geometry = np.array(self._fixed["TemplateData"][0], dtype=np.float64)
        self._precision = int(atlas.desc.precision)

    def GetFixedEffects(self):
        return _copy_effects(self._fixed)

    def SetFixedEffects(self, fixed):
        template_memory = self._fixed["TemplateData"]
        cp_memory = self._fixed["ControlPoints"]
        self._fixed = _copy_effects(fixed)
        if self.freeze_template:
            self._fixed["TemplateData"] = template_memory
        else:
            self._geometry = np.array(template_memory[0], dtype=np.float64)
        if self.freeze_control_points:
            self._fixed["ControlPoints"] = cp_memory

    def _momenta(self, ind_rer):
        return np.stack([np.asarray(a, dtype=np.float64) for a in ind_rer["Momenta"]])

    def UpdateFixedEffectsAndComputeCompleteLogLikelihood(self, pop_rer, ind_rer):
        terms, _ = self.atlas.ComputeLogLikelihood(self._fixed["ControlPoints"],
                                                   self._momenta(ind_rer), self._geometry)
        return np.asarray(terms, dtype=np.float64), False

    def _smooth(self, gl2):
        if self.smoothing_width < 1e-20:
            return gl2
        from .kernel import HIPExactKernel
        tpl = np.asarray(self._fixed["TemplateData"][0], dtype=np.float64)
        k = HIPExactKernel(self.atlas.dim, self.smoothing_width, precision=self._precision)
        k.SetSources(tpl)
        k.SetWeights(gl2)
        return k.Convolve(tpl)

    def ComputeCompleteLogLikelihoodGradient(self, pop_rer, ind_rer):
        _, gcp, gmom, gl2 = self.atlas.ComputeCompleteLogLikelihoodGradient(
            self._fixed["ControlPoints"], self._momenta(ind_rer), self._geometry)
        pop, sq = {}, []
        if not self.freeze_control_points:
            pop["ControlPoints"] = gcp
            sq.append(float(np.sum(gcp * gcp)))
        if not self.freeze_template:
            sob = self._smooth(gl2)
            pop["TemplateData"] = [sob]
            sq.append(float(np.sum(gl2 * sob)))
        ind = {"Momenta": [g for g in gmom]}
        sq.append(float(np.sum(gmom * gmom)))
        return pop, ind, np.asarray(sq)


class HIPMultiObjectAtlasModel:
    """DeterministicAtlas of a template made of several landmark objects (a
    DeformableMultiObject of polylines and triangle surfaces) as the estimators see it, with
    every kernel sum on the GPU: each subject's shooting, flow and the reference's continuous
    adjoint through HIPDiffeos (Diffeos::Update, ::IntegrateAdjointEquations), the data
    terms through the objects' own ComputeMatch / ComputeMatchGradient (HIPOrientedPolyLine,
    HIPNonOrientedPolyLine, HIPOrientedSurfaceMesh, HIPNonOrientedSurfaceMesh,
    HIPOrientedVolumeMesh, HIPPointCloud, HIPLandmark), the RKHS
    regularity and the template smoothing through HIPExactKernel.

    ``objects`` are the template's data-term objects (their kernel widths set),
    ``data_sigma_squared`` one noise variance per object, ``targets`` one list of target
    objects (same types, in object order) per subject.  Fixed effects "ControlPoints"
    (n x d) and "TemplateData" (one point array per object); individual random effects
    "Momenta" (one n x d per subject).

    The terms follow the reference:
    - log-likelihood [-1/2 sum_s sum_i r_si / data_sigma_squared_i, -1/2 sum_s a_s^T K a_s]
      (DeterministicAtlas.cxx:100-151), r_si the match of object i of the deformed template
      with subject s's object i;
    - per subject, each object's match gradient divided by 2 data_sigma_squared_i, the
      objects' points concatenated in object order for the flow and the adjoint
      (AbstractAtlas::ComputeDataTermGradientSubject, AbstractAtlas.cxx:562-616), then the
      regularity (DeterministicAtlas.cxx:236-262);
    - SetFixedEffects' one-call lag of the template geometry, the freeze flags, the
      per-object smoothing with the "TemplateData" fixed effect as sources
      (AbstractAtlas.cxx:52-63, 727-751) and the squared norms [control points, <L2,
      smoothed> summed over objects, momenta] (DeterministicAtlas.cxx:234-238), as in
      HIPAtlasModel.
    The bounding-box test of the reference's Diffeos is not made: out_of_box is always
    False."""

    def __init__(self, objects, data_sigma_squared, targets, fixed_effects, *,
                 deformation_width, number_of_time_points, t0=0.0, tN=1.0,
                 improved_euler=False, scheme="euler", precision="fp64",
                 freeze_template=False, freeze_control_points=False, smoothing_width=0.0,
                 checkpoint_every=0):
        if len(data_sigma_squared) != len(objects):
            raise ValueError("one data sigma squared per object")
        if any(len(t) != len(objects) for t in targets):
            raise ValueError("every subject needs one target per template object")
        self.objects = list(objects)
        self.data_sigma_squared = [float(s) for s in data_sigma_squared]
        self.targets = [list(t) for t in targets]
        self.deformation_width = float(deformation_width)
        self.T = int(number_of_time_points)
        self.t0, self.tN = float(t0), float(tN)
        self.improved_euler = bool(improved_euler)
        self.scheme = scheme
        self.precision = precision
        self.freeze_template = bool(freeze_template)
        self.freeze_control_points = bool(freeze_control_points)
        self.smoothing_width = float(smoothing_width)
        self.checkpoint_every = int(checkpoint_every)
        self._fixed = _copy_effects(fixed_effects)
        self._geometry = [np.array(v, dtype=np.float64) for v in self._fixed["TemplateData"]]
        self._counts = [o.GetNumberOfPoints() for o in self.objects]
        if [len(v) for v in self._geometry] != self._counts:
            raise ValueError("TemplateData must hold each object's points")
        self.dim = self._fixed["ControlPoints"].shape[1]

    def GetFixedEffects(self):
        return _copy_effects(self._fixed)

    def SetFixedEffects(self, fixed):
        template_memory = self._fixed["TemplateData"]
        cp_memory = self._fixed["ControlPoints"]
        self._fixed = _copy_effects(fixed)
        if self.freeze_template:
            self._fixed["TemplateData"] = template_memory
        else:
            self._geometry = [np.array(v, dtype=np.float64) for v in template_memory]
        if self.freeze_control_points:
            self._fixed["ControlPoints"] = cp_memory

    def _diffeos(self, a):
        from .diffeos import HIPDiffeos
        d = HIPDiffeos(self.dim, self.precision)
        d.SetKernelWidth(self.deformation_width)
        d.SetNumberOfTimePoints(self.T)
        d.SetT0(self.t0)
        d.SetTN(self.tN)
        d.SetShootingScheme(self.scheme)
        if self.improved_euler:
            d.UseImprovedEuler()
        else:
            d.UseStandardEuler()
        d.SetStartPositions(self._fixed["ControlPoints"])
        d.SetStartMomentas(np.asarray(a, dtype=np.float64))
        d.SetLandmarkPoints(np.concatenate(self._geometry))
        d.Update()
        return d

    def _deformed(self, d):
        parts = np.split(d.GetDeformedLandmarkPoints(), np.cumsum(self._counts)[:-1])
        return [o.DeformedObject(x) for o, x in zip(self.objects, parts)]

    def _regularity_kernel(self):
        from .kernel import HIPExactKernel
        k = HIPExactKernel(self.dim, self.deformation_width, precision=self.precision)
        k.SetSources(self._fixed["ControlPoints"])
        return k

    def UpdateFixedEffectsAndComputeCompleteLogLikelihood(self, pop_rer, ind_rer):
        momenta = ind_rer["Momenta"]
        if len(momenta) != len(self.targets):
            raise RuntimeError("Number of momenta matrices and residuals mismatch in"
                               " DeterministicAtlas::ComputeLogLikelihood")
        data = 0.0
        for a, target in zip(momenta, self.targets):
            deformed = self._deformed(self._diffeos(a))
            for obj, tgt, s2 in zip(deformed, target, self.data_sigma_squared):
                data -= obj.ComputeMatch(tgt) / s2
        reg = 0.0
        k = self._regularity_kernel()
        cp = self._fixed["ControlPoints"]
        for a in momenta:
            k.SetWeights(a)
            reg -= float(np.sum(k.Convolve(cp) * np.asarray(a)))
        return np.array([0.5 * data, 0.5 * reg]), False

    def _smooth(self, gl2):
        if self.smoothing_width < 1e-20:
            return gl2
        from .kernel import HIPExactKernel
        out = []
        for tpl, g in zip(self._fixed["TemplateData"], gl2):
            k = HIPExactKernel(self.dim, self.smoothing_width, precision=self.precision)
            k.SetSources(tpl)
            k.SetWeights(g)
            out.append(k.Convolve(tpl))
        return out

    def ComputeCompleteLogLikelihoodGradient(self, pop_rer, ind_rer):
        momenta = ind_rer["Momenta"]
        if len(momenta) != len(self.targets):
            raise RuntimeError("Number of Momentas and target multi-object mismatch in"
                               " DeterministicAtlas::ComputeFunctionalGradient")
        cp = self._fixed["ControlPoints"]
        gcp = np.zeros_like(cp)
        gmom = []
        gtpl = np.zeros((sum(self._counts), self.dim))
        for a, target in zip(momenta, self.targets):
            d = self._diffeos(a)
            theta = np.concatenate([obj.ComputeMatchGradient(tgt) / (2.0 * s2) for obj, tgt, s2
                                    in zip(self._deformed(d), target, self.data_sigma_squared)])
            xp, xm, th0 = d.IntegrateAdjointEquations(theta, self.checkpoint_every)
            gcp -= xp
            gmom.append(-xm)
            gtpl -= th0
        k = self._regularity_kernel()
        for s, a in enumerate(momenta):
            k.SetWeights(a)
            gmom[s] -= k.Convolve(cp)
            gcp -= k.ConvolveGradient(cp, np.asarray(a, dtype=np.float64))
        pop, sq = {}, []
        if not self.freeze_control_points:
            pop["ControlPoints"] = gcp
            sq.append(float(np.sum(gcp * gcp)))
        if not self.freeze_template:
            gl2 = np.split(gtpl, np.cumsum(self._counts)[:-1])
            sob = self._smooth(gl2)
            pop["TemplateData"] = sob
            sq.append(sum(float(np.sum(a * b)) for a, b in zip(gl2, sob)))
        sq.append(float(sum(np.sum(g * g) for g in gmom)))
        return pop, {"Momenta": gmom}, np.asarray(sq)
