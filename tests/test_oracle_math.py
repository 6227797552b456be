"""The oracle against mathematics, independent of the reference's outputs.

Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand) fails it:

* edge integrals P_k^m in every regime vs adaptive quadrature of the
  definition (primitives.py:431-504);
* closed-form moment tables vs Gauss quadrature of M_n^(a,b) = int xi^a xi^b
  R^-n (moments.py:171-363), near and at 6 L;
* Hooke consistency: C : sym grad(U) by finite differences == H, for the
  direct and the image halves (verify_hooke_consistency.py, 1e-7);
* eigenstress weight E == Gauss of the blob 15 eps^4 / (8 pi R^7);
* the half-space field is traction-free on z = 0 (verify_image_kernel.py
  TOL_FREE_SURFACE 2e-3);
* the Q-family at the budget-law order vs 4x that order (TOL_BUDGET 1e-7);
* Boussinesq: the surface is traction-free, the load's vertical force is -1
  on every depth plane, and Love's disk equals the point load integrated over
  the disk (fold.py:132-210).
"""
import numpy as np
import pytest
from scipy import integrate

from conftest import relmax
from oracle import edges, fold, fullspace, image, moments
from oracle.geometry import frame_of, gauss_triangle

TRI = np.array([[0.37, -0.81, -1.44], [1.92, 0.11, -2.63], [-0.25, 1.57, -0.78]])
MU, LAM = 1.0, 1.5
NU = 0.5 * LAM / (LAM + MU)


@pytest.mark.parametrize("regime", ["closed", "small", "large", "large_neg", "mixed"])
def test_edge_integrals_vs_quadrature(regime):
    rng = np.random.default_rng(3)
    spec = {-1: 0, 1: 2, 3: 4, 5: 5, 7: 6}
    for _ in range(12):
        rho = rng.uniform(0.2, 2.0)
        if regime == "small":
            ua, ub = np.sort(rng.uniform(-0.5, 0.5, 2)) * rho
        elif regime == "large":
            ua = rng.uniform(2, 6) * rho
            ub = ua + rng.uniform(0.05, 3) * rho
        elif regime == "large_neg":
            ub = -rng.uniform(2, 6) * rho
            ua = ub - rng.uniform(0.05, 3) * rho
        elif regime == "mixed":
            ua, ub = -rng.uniform(0.6, 3) * rho, rng.uniform(0.6, 3) * rho
        else:
            ua = rng.uniform(0.6, 1.5) * rho
            ub = ua + rng.uniform(0.5, 2) * rho
        tab = edges.edge_table(np.array([ua]), np.array([ub]), np.array([rho * rho]), spec)
        for m, kmax in spec.items():
            for k in range(kmax + 1):
                want, _ = integrate.quad(lambda u: u ** k * (u * u + rho * rho) ** (-m / 2.0), ua, ub,
                                         epsabs=0, epsrel=1e-13, limit=200)
                got = tab[m][0, k]
                assert abs(got - want) <= 1e-10 * max(abs(want), 1e-300) + 1e-14, (regime, m, k, got, want)


@pytest.mark.parametrize("where", ["near", "mid"])
def test_closed_form_moments_vs_gauss(where):
    fr = frame_of(TRI)
    rng = np.random.default_rng(4)
    if where == "near":
        inplane = rng.normal(0, 0.5, (6, 2))
        obs = fr.centroid + inplane[:, :1] * fr.e1 + inplane[:, 1:] * fr.e2 + 0.4 * fr.nhat
    else:
        d = rng.normal(0, 1, (6, 3))
        obs = fr.centroid + 6.0 * fr.L * d / np.linalg.norm(d, axis=1)[:, None]
    eps = 0.1
    need = {3: 2, 5: 4, 7: 5, 9: 5}
    z, X, M = moments.raw_moments(fr, obs, eps, need)
    x1, x2, w = gauss_triangle(120)
    lam = np.stack([1 - x1 - x2, x1, x2], axis=1)
    eta = lam @ fr.p
    wq = w * 2 * fr.area
    for n, d in need.items():
        for a in range(d + 1):
            for b in range(d + 1 - a):
                xi = eta[None] - X[:, None]
                R2 = (xi ** 2).sum(axis=2) + (z ** 2 + eps ** 2)[:, None]
                want = (wq[None] * xi[..., 0] ** a * xi[..., 1] ** b * R2 ** (-n / 2)).sum(axis=1)
                tol = 1e-9 if where == "near" else 1e-6
                assert relmax(M[n][:, a, b], want) < tol, (n, a, b)


def _hooke_fd(Ufun, obs, h=1e-5):
    """C : sym grad(u) by central differences of u(x) = U[n, k, i, slip]."""
    grads = []
    for ax in range(3):
        e = np.zeros(3)
        e[ax] = h
        grads.append((Ufun(obs + e) - Ufun(obs - e)) / (2 * h))      # (N, K, i, slip) per axis
    G = np.stack(grads, axis=-1)                                       # (N, K, i, slip, ax)
    eps_ = 0.5 * (G + np.swapaxes(G, 2, 4))
    tr = eps_[:, :, 0, :, 0] + eps_[:, :, 1, :, 1] + eps_[:, :, 2, :, 2]
    sig = 2 * MU * eps_
    for a in range(3):
        sig[:, :, a, :, a] += LAM * tr
    return np.moveaxis(sig, 3, 4)                                      # (N, K, a, b, slip)


@pytest.mark.parametrize("half", ["direct", "image"])
def test_hooke_consistency(half):
    rng = np.random.default_rng(5)
    obs = TRI.mean(axis=0) + rng.normal(0, 1.5, (5, 3))
    obs[:, 2] = np.minimum(obs[:, 2], -0.3)
    if half == "direct":
        U = lambda x: fullspace.nodal(x, TRI, 1, MU, NU, 0.1, ("U",))["U"]
        H = fullspace.nodal(obs, TRI, 1, MU, NU, 0.1, ("H",))["H"]
    else:
        U = lambda x: image.image_total(x, TRI, 1, MU, LAM, 0.1, ("U",), n_quad=24)["U"]
        H = image.image_total(obs, TRI, 1, MU, LAM, 0.1, ("H",), n_quad=24)["H"]
    assert relmax(_hooke_fd(U, obs), H) < 1e-7


def test_eigen_weight_is_the_blob_integral():
    fr = frame_of(TRI)
    eps = 0.3
    obs = fr.centroid + np.array([[0.1, 0.2, 0.0], [0.5, -0.3, 0.2]]) + 0.25 * fr.nhat
    E = fullspace.nodal(obs, TRI, 0, MU, NU, eps, ("E",))["E"][:, 0]
    x1, x2, w = gauss_triangle(160)
    y = np.stack([1 - x1 - x2, x1, x2], axis=1) @ TRI
    R2 = ((obs[:, None] - y[None]) ** 2).sum(axis=2) + eps ** 2
    want = (w * 2 * fr.area * 15 * eps ** 4 / (8 * np.pi) * R2 ** -3.5).sum(axis=1)
    assert relmax(E, want) < 1e-10


def test_free_surface_is_traction_free():
    """Direct + image total stress of a buried source has sigma . z = 0 on z = 0
    (away from the source's footprint)."""
    tri = TRI - np.array([0, 0, 1.0])
    rng = np.random.default_rng(6)
    obs = np.column_stack([rng.uniform(-4, 4, 12), rng.uniform(-4, 4, 12), np.zeros(12)])
    H = (fullspace.nodal(obs, tri, 0, MU, NU, 0.05, ("H",))["H"]
         + image.image_total(obs, tri, 0, MU, LAM, 0.05, ("H",))["H"])
    trac = H[:, 0, :, 2, :]                      # sigma_{i z} per slip
    assert np.max(np.abs(trac)) / np.max(np.abs(H)) < 2e-3


def test_q_family_converged_at_law_order():
    tri = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, -0.9]])
    obs = np.array([[0.1, 0.05, -0.2], [0.3, 0.4, -0.6], [2.0, -1.0, -1.0]])
    law = image.q_orders(obs, tri * np.array([1, 1, -1.0]), 0.1, frame_of(tri).L, 0)
    a = image.image_q(obs, tri, 0, MU, LAM, 0.1, ("H",))["H"]
    for i, o in enumerate(obs):
        b = image.image_q(o[None], tri, 0, MU, LAM, 0.1, ("H",), n_quad=min(4 * int(law[i]), 192))["H"]
        assert relmax(a[i:i + 1], b) < 1e-7


def test_boussinesq_point_load():
    # traction-free surface away from the load
    xy = np.array([0.7, -0.4])
    S = fold.point_stress(np.array([xy[0]]), np.array([xy[1]]), 0.0, NU)[0]
    assert np.max(np.abs(S[2])) < 1e-14
    # vertical force on a depth plane: int sigma_zz dA = -1 (unit compressive load)
    zeta = 0.5
    r = np.linspace(0, 400, 400001)
    szz = fold.point_stress(r, np.zeros_like(r), zeta, NU)[:, 2, 2]
    force = integrate.trapezoid(szz * 2 * np.pi * r, r)
    assert abs(force + 1.0) < 1e-4
    # Love's disk = point load integrated over the disk, on axis
    a = 1.3
    rr = np.linspace(0, a, 4001)
    Srr = fold.point_stress(rr, np.zeros_like(rr), zeta, NU)
    # azimuthal average of the in-plane tensor on the axis: (sxx + syy) / 2
    srr = integrate.trapezoid(0.5 * (Srr[:, 0, 0] + Srr[:, 1, 1]) * 2 * np.pi * rr, rr)
    szz_disk = integrate.trapezoid(Srr[:, 2, 2] * 2 * np.pi * rr, rr)
    love = fold.disk_stress(a, zeta, NU)
    assert abs(szz_disk - love[2, 2]) < 1e-6
    assert abs(srr - love[0, 0]) < 1e-6
