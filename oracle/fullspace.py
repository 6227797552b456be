"""Full-space (Kelvin-Cortez) nodal kernels U, H, E of one triangle.

Restates src/mhs/fullspace/kernels.py:63-213: the lift of weighted moments to
Cartesian tensor moments, and the slip-source contractions

    U  (N, K, 3, 3)     displacement i per unit Cartesian slip j
    H  (N, K, 3, 3, 3)  TOTAL stress (m, l) per unit slip j
    E  (N, K)           blob weight, eigenstress = E * C:sym(slip (x) n)

with the traction-operator pairing and C1 = 1/(16 pi mu (1 - nu)).
"""
from __future__ import annotations

import numpy as np

from .geometry import frame_of
from .moments import kernel_degrees, weighted_tables


def lift_groups(basis, rank: int) -> dict:
    """{(a, b): sum of basis outer products with a e1's and b e2's}
    (kernels.py:63-87)."""
    geom = {(0, 0): np.array(1.0)}
    for _ in range(rank):
        nxt = {}
        for (a, b), g in geom.items():
            for s, v in enumerate(basis):
                key = (a + (s == 0), b + (s == 1))
                o = np.multiply.outer(g, v)
                nxt[key] = nxt[key] + o if key in nxt else o
        geom = nxt
    return geom


def lift(W, z, fr, rank: int, n: int) -> np.ndarray:
    """int N_k d_i1..d_ir / R^n as (N, K, 3^rank) (kernels.py:90-123), with
    d = x - y = -xi1 e1 - xi2 e2 + z nhat."""
    Wn = W[n]
    N, K = Wn.shape[:2]
    out = np.zeros((N, K) + (3,) * rank)
    zp = {}
    for (a, b), g in lift_groups((fr.e1, fr.e2, fr.nhat), rank).items():
        c = rank - a - b
        if c not in zp:
            zp[c] = z ** c
        scal = ((-1.0) ** (a + b)) * zp[c][:, None] * Wn[:, :, a, b]
        out += scal.reshape((N, K) + (1,) * rank) * g
    return out


def nodal(obs, tri, order, mu, nu, eps, want=("U", "H", "E"), far_field="hybrid"):
    """kernels.py:126-213 for the slip source."""
    fr = frame_of(tri)
    obs = np.asarray(obs, float).reshape(-1, 3)
    W, z, X = weighted_tables(fr, obs, eps, order, kernel_degrees(0, want), far_field)
    lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
    C1 = 1.0 / (16.0 * np.pi * mu * (1.0 - nu))
    c34 = 3.0 - 4.0 * nu
    e2 = eps * eps
    n = fr.nhat
    eye = np.eye(3)
    out = {}
    if "U" in want:
        V3 = lift(W, z, fr, 1, 3)
        V5 = lift(W, z, fr, 1, 5)
        T35 = lift(W, z, fr, 3, 5)
        G1 = (-c34 * np.einsum("ij,nkm->nkijm", eye, V3)
              + np.einsum("im,nkj->nkijm", eye, V3)
              + np.einsum("jm,nki->nkijm", eye, V3)
              - 3.0 * T35
              - 6.0 * (1.0 - nu) * e2 * np.einsum("ij,nkm->nkijm", eye, V5))
        G1 *= C1
        out["U"] = -(mu * np.einsum("m,nkijm->nkij", n, G1)
                     + lam * np.einsum("j,nkimm->nkij", n, G1)
                     + mu * np.einsum("m,nkimj->nkij", n, G1))
    if "H" in want:
        I3 = W[3][:, :, 0, 0]
        I5 = W[5][:, :, 0, 0]
        T25 = lift(W, z, fr, 2, 5)
        T27 = lift(W, z, fr, 2, 7)
        T47 = lift(W, z, fr, 4, 7)
        D2 = np.zeros(T47.shape)
        D2 += -c34 * (np.einsum("rp,sq,nk->nkrpsq", eye, eye, I3)
                      - 3.0 * np.einsum("rp,nksq->nkrpsq", eye, T25))
        D2 += (np.einsum("rs,pq,nk->nkrpsq", eye, eye, I3)
               - 3.0 * np.einsum("rs,nkpq->nkrpsq", eye, T25))
        D2 += (np.einsum("ps,rq,nk->nkrpsq", eye, eye, I3)
               - 3.0 * np.einsum("ps,nkrq->nkrpsq", eye, T25))
        D2 += -3.0 * (np.einsum("rq,nkps->nkrpsq", eye, T25)
                      + np.einsum("pq,nkrs->nkrpsq", eye, T25)
                      + np.einsum("sq,nkrp->nkrpsq", eye, T25))
        D2 += 15.0 * T47
        D2 += 2.0 * (1.0 - nu) * e2 * (-3.0 * np.einsum("rp,sq,nk->nkrpsq", eye, eye, I5)
                                      + 15.0 * np.einsum("rp,nksq->nkrpsq", eye, T27))
        D2 *= C1
        trD = np.einsum("nkrpsp->nkrs", D2)
        B = (lam * np.einsum("j,nkrs->nkrsj", n, trD)
             + mu * np.einsum("q,nkrjsq->nkrsj", n, D2)
             + mu * np.einsum("p,nkrpsj->nkrsj", n, D2))
        trB = np.einsum("nkrrj->nkj", B)
        out["H"] = -(lam * np.einsum("ml,nkj->nkmlj", eye, trB)
                     + mu * (B + np.swapaxes(B, 2, 3)))
    if "E" in want:
        out["E"] = (15.0 * eps ** 4 / (8.0 * np.pi)) * W[7][:, :, 0, 0]
    return out


def blocks(obs, tri, mu, lam, eps, want, order):
    """The (mu, lam) seam (assemble.py:58-66, 96-105): nu = lam / (2 (lam+mu))."""
    return nodal(np.asarray(obs, float), np.asarray(tri, float), order,
                 float(mu), 0.5 * lam / (lam + mu), float(eps), want=tuple(want))
