"""The FP64 oracle against the reference's own outputs (golden vectors written
by tests/golden/make_golden.py in the build container)."""
import numpy as np
import pytest

from conftest import GOLDEN, relmax
from oracle import api as O
from oracle import fold as F
from oracle import image as I

TOL = 1e-12


def _mat(g):
    return float(g["mu"]), float(g["lam"])


@pytest.mark.parametrize("p", [0, 1, 2])
def test_buried_all_forms(small, p):
    g = small
    mu, lam = _mat(g)
    tb, eb, ob = g["tb"], g["eb"], g["ob"]
    assert relmax(O.disp_matrix(ob, tb, mu, lam, eb, order=p), g[f"disp_b_p{p}"]) < TOL
    sig, eig = O.stress_matrix(ob, tb, mu, lam, eb, parts=True, order=p)
    assert relmax(sig, g[f"stress_b_p{p}"]) < TOL
    assert relmax(eig, g[f"eig_b_p{p}"]) < TOL
    assert relmax(O.traction_matrix(ob, g["nrm_b"], tb, mu, lam, eb, order=p), g[f"trac_b_p{p}"]) < TOL
    assert relmax(O.interaction_matrix(tb, mu, lam, eb, order=p), g[f"inter_b_p{p}"]) < TOL


def test_buried_variants(small):
    g = small
    mu, lam = _mat(g)
    tb, eb, ob = g["tb"], g["eb"], g["ob"]
    assert relmax(O.total_stress_matrix(ob, tb, mu, lam, eb), g["total_b_p0"]) < TOL
    assert relmax(O.elastic_strain_matrix(ob, tb, mu, lam, eb), g["strain_b_p0"]) < TOL
    assert relmax(O.traction_matrix(ob, g["nrm_b"], tb, mu, lam, eb, subtract_eigenstress=False),
                  g["trac_b_noeig"]) < TOL
    assert relmax(O.interaction_matrix(tb, mu, lam, eb, receiver="strike", source="strike"),
                  g["inter_b_ss"]) < TOL
    assert relmax(O.interaction_matrix(tb, mu, lam, eb, order=1, rows=(5, 17), receiver=("dip", "normal"),
                                       source="tensile"), g["inter_b_rows"]) < TOL
    assert relmax(O.interaction_matrix(tb, mu, lam, eb, obs_tris=tb[:3], shrink=0.2, order=1),
                  g["inter_b_obs"]) < TOL
    assert relmax(O.interaction_matrix(tb, mu, lam, eb, order=2, layout="continuous"), g["inter_b_cont"]) < TOL
    assert relmax(O.disp_matrix(ob, tb, mu, lam, eb, order=2, layout="continuous"), g["disp_b_cont"]) < TOL


def test_surface_fold_p0(small):
    """A surface-reaching mesh: the fold (mirror + Boussinesq) at P0."""
    g = small
    mu, lam = _mat(g)
    ts, es, os_ = g["ts"], float(g["es"]), g["os"]
    assert relmax(O.interaction_matrix(ts, mu, lam, es), g["inter_s_p0"]) < TOL
    assert relmax(O.disp_matrix(os_, ts, mu, lam, es, fold=False), g["disp_s_nofold"]) < TOL


def test_q_orders():
    g = np.load(GOLDEN / "qorders.npz")
    obs, tris = g["obs"], g["tris"]
    from oracle.geometry import frame_of
    rows = []
    for s in range(tris.shape[0]):
        L = frame_of(tris[s]).L
        for p in (0, 2):
            for eps in (0.01, 0.1):
                rows.append(I.q_orders(obs, tris[s] * np.array([1, 1, -1.0]), eps, L, p))
    assert np.array_equal(np.array(rows), g["orders"])


def test_fold_grid_and_image():
    g = dict(np.load(GOLDEN / "fold.npz"))
    tris, obs = g["tris"], g["obs"]
    eps = 0.1
    for s in g["s_idx"]:
        c, r = F.crease(tris[s], eps)
        assert np.allclose([c[0], c[1], r], g[f"crease_{s}"], rtol=0, atol=1e-15)
        diam = max(float(np.linalg.norm(tris[s][i] - tris[s][j])) for i in range(3) for j in range(3))
        for o in range(2):
            pts, wts, R1 = F.polar_grid(obs[o, :2], max(0.0, -obs[o, 2]), c, r, eps, diam)
            assert np.array_equal(pts, g[f"grid_{s}_{o}_pts"])
            assert np.array_equal(wts, g[f"grid_{s}_{o}_wts"])
            assert R1 == float(g[f"grid_{s}_{o}_R1"])
        img = F.image(obs, tris[s], 30.0, 45.0, eps, 0, ("U", "H"))
        assert relmax(img["U"], g[f"img_{s}_p0_U"]) < TOL
        assert relmax(img["H"], g[f"img_{s}_p0_H"]) < TOL
