"""The assembly seam's argument checks (CPU, before any device work): the
reference's assemble_* signatures (src/mhs/kernels/assemble.py:186-366) with
wrong shapes, ids or out are refused with a message naming the argument."""
import inspect

import numpy as np
import pytest

from mhs_amd import seam


def _mesh(n=4):
    rng = np.random.default_rng(2)
    return rng.normal(size=(n, 3, 3)) - [0.0, 0.0, 6.0]


def test_signatures_are_the_references():
    """Positional order and keyword defaults of assemble.py:186-366."""
    want = {
        "assemble_halfspace_disp": "(obs, tris, mu, lam, eps_arr, out, order=0, ids=None, fold=True)",
        "assemble_halfspace_total_stress": "(obs, tris, mu, lam, eps_arr, out, order=0, ids=None, fold=True)",
        "assemble_halfspace_traction": "(obs, nrm, tris, mu, lam, eps_arr, out, subtract_eigenstress=True, "
                                       "order=0, ids=None, fold=True)",
        "assemble_interaction": "(obs, nrm, rec_basis, tris, src_basis, mu, lam, eps_arr, out, "
                                "subtract_eigenstress=True, order=0, ids=None, fold=True)",
        "assemble_eigenstress": "(obs, tris, mu, lam, eps_arr, out, order=0, ids=None, fold=True)",
    }
    for name, sig in want.items():
        assert str(inspect.signature(getattr(seam, name))) == sig, name


def test_refusals_before_device_work():
    tris = _mesh()
    obs = np.zeros((5, 3)) - [0, 0, 1.0]
    eps = np.full(4, 0.1)
    with pytest.raises(ValueError, match=r"out has shape \(5, 3, 3, 3\), expected \(5, 3, 4, 3\)"):
        seam.assemble_halfspace_disp(obs, tris, 30.0, 30.0, eps, np.zeros((5, 3, 3, 3)))
    with pytest.raises(ValueError, match="eps_arr must have one value per source triangle"):
        seam.assemble_halfspace_disp(obs, tris, 30.0, 30.0, eps[:3], np.zeros((5, 3, 4, 3)))
    with pytest.raises(ValueError, match="strictly positive"):
        seam.assemble_halfspace_disp(obs, tris, 30.0, 30.0, 0.0 * eps, np.zeros((5, 3, 4, 3)))
    with pytest.raises(ValueError, match="order must be 0, 1 or 2"):
        seam.assemble_eigenstress(obs, tris, 30.0, 30.0, eps, np.zeros((5, 3, 3, 4, 3)), order=3)
    with pytest.raises(ValueError, match=r"obs must have shape"):
        seam.assemble_halfspace_disp(obs[:, :2], tris, 30.0, 30.0, eps, np.zeros((5, 3, 4, 3)))
    with pytest.raises(ValueError, match=r"nrm must have shape \(5, 3\)"):
        seam.assemble_halfspace_traction(obs, obs[:4], tris, 30.0, 30.0, eps, np.zeros((5, 3, 4, 3)))
    with pytest.raises(ValueError, match="float64 or float32"):
        seam.assemble_halfspace_disp(obs, tris, 30.0, 30.0, eps, np.zeros((5, 3, 4, 3), np.int64))
    with pytest.raises(TypeError, match="numpy.ndarray"):
        seam.assemble_halfspace_disp(obs, tris, 30.0, 30.0, eps, [0.0])
    with pytest.raises(ValueError, match=r"ids must have shape \(4, 3\)"):
        seam.assemble_halfspace_disp(obs, tris, 30.0, 30.0, eps, np.zeros((5, 3, 7, 3)), order=1,
                                     ids=np.zeros((4, 2), int))
    with pytest.raises(ValueError, match=r"ids must name columns in \[0, 7\)"):
        seam.assemble_halfspace_disp(obs, tris, 30.0, 30.0, eps, np.zeros((5, 3, 7, 3)), order=1,
                                     ids=np.full((4, 3), 7))
    rb = np.zeros((5, 2, 3))
    with pytest.raises(ValueError, match=r"src_basis must have shape \(4, n_b, 3\)"):
        seam.assemble_interaction(obs, obs, rb, tris, np.zeros((3, 1, 3)), 30.0, 30.0, eps, np.zeros((5, 4, 2, 1)))
    with pytest.raises(ValueError, match=r"rec_basis must have shape \(5, n_a, 3\)"):
        seam.assemble_interaction(obs, obs, rb[:4], tris, np.zeros((4, 1, 3)), 30.0, 30.0, eps,
                                  np.zeros((5, 4, 2, 1)))
    with pytest.raises(ValueError, match=r"out has shape \(5, 4, 1, 1\), expected \(5, 4, 2, 1\)"):
        seam.assemble_interaction(obs, obs, rb, tris, np.zeros((4, 1, 3)), 30.0, 30.0, eps, np.zeros((5, 4, 1, 1)))
