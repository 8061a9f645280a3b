"""ORACLE — test infrastructure only.

FP64 CPU restatement of Deformetrica's PointCloud and Landmark data terms on top of the
oracle ExactKernel:

  PointCloud  weights  one scalar per point (the VTK point scalars, else 1;
                       PointCloud.cxx:203-226)
              norm     |S|^2 = sum_i w_i (K w)(x_i) (UpdateSelfNorm, :231-250)
              match    |T|^2 + |S|^2 - sum_j (K w_S)(xt_j) * 2 wt_j (ComputeMatch, :108-148)
              gradient 2 (ConvolveGradient(x, w) from (x, w) + ConvolveGradient(x, w) from
                       (xt, wt)) (ComputeMatchGradient, :152-194): the cross term enters with
                       +, so the formula is d/dx |S + T|^2 rather than the derivative of the
                       match
  Landmark    match    |X - Y|_F^2 (Landmark.cxx:287-304)
              gradient 2 (X - Y) (:306-322)

Pinned by tests/test_points.py: the point-cloud match against a brute-force double sum over
point pairs, the gradient against central differences of the brute-force |S + T|^2 (and,
with the cross term's sign turned, of the match itself), the landmark match against
numpy's Frobenius norm and its gradient against central differences.
"""
from __future__ import annotations

import numpy as np

from .oracle import ExactKernel


def _kernel(Y, w, width):
    k = ExactKernel(Y.shape[1], width)
    k.SetSources(np.asarray(Y, dtype=np.float64))
    k.SetWeights(np.asarray(w, dtype=np.float64).reshape(-1, 1))
    return k


def pointcloud_norm2(X, w, width):
    X = np.asarray(X, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    if X.shape[0] == 0:
        return 0.0
    return float(np.sum(_kernel(X, w, width).Convolve(X)[:, 0] * w))


def pointcloud_match(Xs, ws, Xt, wt, width):
    Xt = np.asarray(Xt, dtype=np.float64)
    wt = np.asarray(wt, dtype=np.float64)
    match = pointcloud_norm2(Xt, wt, width) + pointcloud_norm2(Xs, ws, width)
    if Xt.shape[0] and np.asarray(Xs).shape[0]:
        sdot = _kernel(Xs, ws, width).Convolve(Xt)[:, 0]
        for i in range(Xt.shape[0]):
            match -= sdot[i] * (2.0 * wt[i])
    return match


def cross_gradient(Xs, ws, Y, wy, width):
    """ConvolveGradient(Xs, ws) with (Y, wy) as sources: sum_j grad_1 K(x_i, y_j) w_i wy_j."""
    Xs = np.asarray(Xs, dtype=np.float64)
    if Xs.shape[0] == 0 or np.asarray(Y).shape[0] == 0:
        return np.zeros_like(Xs)
    return _kernel(Y, wy, width).ConvolveGradient(Xs, np.asarray(ws, dtype=np.float64).reshape(-1, 1))


def pointcloud_match_gradient(Xs, ws, Xt, wt, width):
    g_ss = cross_gradient(Xs, ws, Xs, ws, width)
    g_st = cross_gradient(Xs, ws, Xt, wt, width)
    return (g_ss + g_st) * 2.0


def brute_force_pointcloud(Xs, ws, Xt, wt, width, sign=-1.0):
    """|S + sign T|^2 of the two weighted Dirac sums as an explicit double sum over points."""
    c = np.concatenate([np.asarray(Xs, dtype=np.float64), np.asarray(Xt, dtype=np.float64)])
    w = np.concatenate([np.asarray(ws, dtype=np.float64), sign * np.asarray(wt, dtype=np.float64)])
    r2 = np.sum((c[:, None, :] - c[None, :, :]) ** 2, axis=-1)
    return float(np.sum(np.exp(-r2 / width ** 2) * np.outer(w, w)))


def landmark_match(X, Y):
    d = np.asarray(X, dtype=np.float64) - np.asarray(Y, dtype=np.float64)
    return float(np.sum(d * d))


def landmark_match_gradient(X, Y):
    return (np.asarray(X, dtype=np.float64) - np.asarray(Y, dtype=np.float64)) * 2.0
