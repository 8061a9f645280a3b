"""The inputs of three of the reference's examples (starfish, bundles_cortico, grid) as the
loaders read them, on the CPU; their runs are checked on the GPU against the oracle."""
import numpy as np

import functional_cases as fc


def test_starfish_inputs():
    template, objects, targets = fc.starfish()
    assert template[0].shape == (194, 2) and objects[0][1].shape == (194, 2)
    (V, E), = targets[0]
    assert V.shape == (190, 2) and E.shape == (190, 2)
    fixed, _ = fc.multi_initial_effects(template, objects, fc.STARFISH_PARAMS["deformation_width"], 1)
    assert fixed["ControlPoints"].shape == (25, 2)


def test_bundles_cortico_inputs():
    template, objects, targets = fc.bundles_cortico()
    assert [v.shape for v in template] == [(1420, 3), (642, 3)]
    assert [np.asarray(o[1]).shape for o in objects] == [(710, 2), (1280, 3)]
    (bv, be), (pv, pf) = targets[0]
    assert bv.shape == (7161, 3) and be.shape == (6663, 2)
    assert pv.shape == (1102, 3) and pf.shape == (2204, 3)


def test_grid_inputs():
    """The control points are the grid's first 24 points, which the VTK file holds in single
    precision."""
    cp, mom, X = fc.grid()
    assert cp.shape == mom.shape == (24, 2) and X.shape == (783, 2)
    np.testing.assert_array_equal(X[:24], cp.astype(np.float32).astype(np.float64))
