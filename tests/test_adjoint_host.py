"""Host-side pieces of the adjoint problems (no GPU): objective weights and parameter maps."""
import math

import numpy as np

from dugks_amd import adjoint, cases


def test_bottom_wall_weight_selects_the_bottom_faces():
    c = cases.cavity_case(5, "NC38", 7, 4.0, 0.1)
    w = adjoint.bottom_wall_q_weight(c)
    a = c.arrays()
    rows = np.nonzero(np.any(w != 0, axis=1))[0]
    assert len(rows) == 5 and np.all(rows >= a["nIF"])
    assert np.allclose(a["Cf"][rows, 1], 0.0)
    # outward normal of the bottom wall points to -y; the weights are the face area vectors
    assert np.allclose(w[rows], a["Sf"][rows]) and np.all(w[rows, 1] < 0)
    # the bottom wall sits in the middle third of fixedWalls, where cavity_case puts wall_T
    fw = adjoint.face_ids_of_patch(c, "fixedWalls")
    assert np.array_equal(fw[5:10], rows)


def test_drag_weight_and_seeded_patches():
    kw = dict(nc=4, nr=2, n_up=2, n_down=3, n_side=2)
    c = cases.sphere_case(kw, "GH", 6, 1.0, 0.1)
    w = adjoint.drag_mom_weight(c)
    ids = adjoint.face_ids_of_patch(c, "sphere")
    U = c.meta["U_inf"]
    assert np.allclose(w[ids, 0], 1.0 / (0.5 * U * U * math.pi / 4.0))
    assert np.count_nonzero(w) == len(ids)
    lab = adjoint.seeded_surface_patches(c, 8)
    assert lab.shape == (len(ids),) and lab.min() >= 0 and lab.max() < 8
    assert np.array_equal(lab, adjoint.seeded_surface_patches(c, 8))
    prob = adjoint.SphereDrag(mesh_kw=kw, n_dv=6, n_patch=8)
    th = prob.nominal()
    th[1:] = np.arange(8) + 1.0
    cc = prob.case(th)
    bc = {b.name: b for b in cc.bcs}["sphere"]
    assert np.array_equal(bc.T, th[1:][prob._labels])
