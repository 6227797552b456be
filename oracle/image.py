"""Mindlin image correction: closed-form R-family plus quadrature Q-family.

Restates src/mhs/kernels/image.py:95-514.  The monomial records come from
``mhs_amd/data/mindlin_table.json`` (this repository's own sympy derivation,
checked record-for-record against the reference's generated table by
tests/test_table.py): every R-family record is

    coeff(z, nu) / (pi mu (1 - nu)) * D1^a D2^b D3^c / R2^n

with D = x - y' (y' the reflected source point) and z the observer's absolute
depth; every Q-family record additionally carries Q^-q, Q = R2 - D3.
"""
from __future__ import annotations

import functools
import json
import pathlib

import numpy as np

from .fullspace import lift
from .geometry import frame_of, gauss_triangle, nodal_at, unit_normal
from .moments import weighted_tables

TABLE_PATH = pathlib.Path(__file__).resolve().parents[1] / "mhs_amd" / "data" / "mindlin_table.json"
REFLECT = np.array([1.0, 1.0, -1.0])
IMAGE_Q_BUDGET_C = 10.0          # defaults.py:60
IMAGE_Q_GAUSS_MIN = 8            # defaults.py:62
IMAGE_Q_GAUSS_MAX = 192          # defaults.py:66


@functools.lru_cache(maxsize=1)
def table():
    return json.loads(TABLE_PATH.read_text())


@functools.lru_cache(maxsize=1)
def degrees():
    """Constant-slip degree per n from the R records (image.py:73-90)."""
    t = table()
    need = {}
    for rec in t["r_dg"]:
        a, b, c, n = rec[3:7]
        need[n] = max(need.get(n, -1), a + b + c)
    for rec in t["r_ddg"]:
        a, b, c, n = rec[4:8]
        need[n] = max(need.get(n, -1), a + b + c)
    return dict(sorted(need.items()))


def _poly(terms, z, nu):
    out = np.zeros_like(z)
    for pz, pnu, num, den in terms:
        out += (num / den) * (z ** pz) * (nu ** pnu)
    return out


def contract(DG, DDG, nrm, mu, lam, want):
    """Traction-operator contraction shared by R and Q (image.py:107-162)."""
    out = {}
    if "U" in want:
        tr = DG[:, :, :, 0, 0] + DG[:, :, :, 1, 1] + DG[:, :, :, 2, 2]
        U = np.zeros(DG.shape[:2] + (3, 3))
        for i in range(3):
            for k in range(3):
                U[:, :, i, k] = (mu * sum(nrm[l] * DG[:, :, i, k, l] for l in range(3))
                                 + mu * sum(nrm[l] * DG[:, :, i, l, k] for l in range(3))
                                 + lam * nrm[k] * tr[:, :, i])
        out["U"] = U
    if "H" in want:
        S = np.zeros(DDG.shape[:2] + (3, 3, 3))
        for i in range(3):
            for p in range(3):
                tr_m = sum(DDG[:, :, i, m, p, m] for m in range(3))
                for k in range(3):
                    S[:, :, i, p, k] = (mu * sum(nrm[l] * DDG[:, :, i, k, p, l] for l in range(3))
                                        + mu * sum(nrm[l] * DDG[:, :, i, l, p, k] for l in range(3))
                                        + lam * nrm[k] * tr_m)
        H = np.zeros(DDG.shape[:2] + (3, 3, 3))
        for k in range(3):
            g = S[:, :, :, :, k]
            e = 0.5 * (g + np.swapaxes(g, -1, -2))
            trace = e[:, :, 0, 0] + e[:, :, 1, 1] + e[:, :, 2, 2]
            for a in range(3):
                for b in range(3):
                    H[:, :, a, b, k] = 2.0 * mu * e[:, :, a, b]
                    if a == b:
                        H[:, :, a, b, k] += lam * trace
        out["H"] = H
    return out


def triangle_distance(obs, tri):
    """Exact point-triangle distance (image.py:170-208)."""
    obs = np.asarray(obs, float).reshape(-1, 3)
    a, b, c = tri[0], tri[1], tri[2]
    ab, ac = b - a, c - a
    nrm = np.cross(ab, ac)
    ap = obs - a
    d20, d21 = ap @ ab, ap @ ac
    d00, d01, d11 = float(ab @ ab), float(ab @ ac), float(ac @ ac)
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    u = 1.0 - v - w
    inside = (u >= 0.0) & (v >= 0.0) & (w >= 0.0)
    out = np.empty(obs.shape[0])
    out[inside] = np.abs(ap[inside] @ nrm) / np.linalg.norm(nrm)
    sel = ~inside
    if np.any(sel):
        best = np.full(int(sel.sum()), np.inf)
        for p0, p1 in ((a, b), (b, c), (c, a)):
            e = p1 - p0
            t = np.clip(((obs[sel] - p0) @ e) / float(e @ e), 0.0, 1.0)
            best = np.minimum(best, np.linalg.norm(obs[sel] - (p0[None] + t[:, None] * e[None]), axis=1))
        out[sel] = best
    return out


def q_orders(obs, tri_img, eps, L, order):
    """Per-observer Q-family Gauss order (image.py:211-254)."""
    delta = triangle_distance(obs, tri_img)
    nq = np.ceil(IMAGE_Q_BUDGET_C * float(L) / np.sqrt(delta * delta + float(eps) ** 2)) + int(order)
    return np.clip(nq, IMAGE_Q_GAUSS_MIN, IMAGE_Q_GAUSS_MAX).astype(int)


@functools.lru_cache(maxsize=8)
def q_basis(nu: float):
    """(monos, C_U (27, n_mono), C_H (81, n_mono)) (image.py:257-297)."""
    t = table()
    monos = {}
    blocks = []
    for key, n_ix in (("q_dg", 3), ("q_ddg", 4)):
        rows = {}
        for rec in t[key]:
            slot = 0
            for ix in rec[:n_ix]:
                slot = slot * 3 + ix
            k = monos.setdefault(tuple(rec[n_ix:n_ix + 5]), len(monos))
            cv = sum((num / den) * nu ** pnu for _pz, pnu, num, den in rec[n_ix + 5])
            rows[(slot, k)] = rows.get((slot, k), 0.0) + cv
        blocks.append((3 ** n_ix, rows))
    n_mono = len(monos)
    mats = []
    for size, rows in blocks:
        C = np.zeros((size, n_mono))
        for (s, k), v in rows.items():
            C[s, k] += v
        mats.append(C)
    return tuple(monos), mats[0], mats[1]


def _q_group(obs, tri, fr, order, nu, scale_c, eps, nq, want, K):
    """One Gauss order for a group of observers (image.py:315-377)."""
    x1, x2, w = gauss_triangle(nq)
    bary = np.stack([1.0 - x1 - x2, x1, x2], axis=1)
    y = bary @ tri
    wN = (w * 2.0 * fr.area)[:, None] * nodal_at(fr, order, bary @ fr.p)
    monos, CU, CH = q_basis(nu)
    d1 = obs[:, 0:1] - y[None, :, 0]
    d2 = obs[:, 1:2] - y[None, :, 1]
    d3 = obs[:, 2:3] + y[None, :, 2]
    r2 = np.sqrt(d1 * d1 + d2 * d2 + d3 * d3 + eps ** 2)
    ir = 1.0 / r2
    iq = 1.0 / (r2 - d3)
    basis = np.empty((len(monos), obs.shape[0], K))
    # power ladders by repeated products (image.py:299-312 caches them too)
    ladders = {}

    def power(name, base, e):
        lad = ladders.setdefault(name, [None, base])
        while len(lad) <= e:
            lad.append(lad[-1] * base)
        return lad[e]

    spatial, radial = {}, {}
    for k, (a, b, c, n, q) in enumerate(monos):
        ds = spatial.get((a, b, c))
        if ds is None:
            ds = None
            for nm, base, e in (("d1", d1, a), ("d2", d2, b), ("d3", d3, c)):
                if e:
                    ds = power(nm, base, e) if ds is None else ds * power(nm, base, e)
            spatial[(a, b, c)] = ds
        rq = radial.get((n, q))
        if rq is None:
            rq = None
            if n > 0:
                rq = power("ir", ir, n)
            elif n < 0:
                rq = power("r", r2, -n)
            if q:
                rq = power("iq", iq, q) if rq is None else rq * power("iq", iq, q)
            radial[(n, q)] = rq
        if ds is None and rq is None:
            t = np.ones_like(r2)
        elif ds is None:
            t = rq
        elif rq is None:
            t = ds
        else:
            t = ds * rq
        basis[k] = t @ wN
    dg = ddg = None
    if "U" in want:
        dg = (np.tensordot(CU, basis, axes=([1], [0])) * scale_c)
        dg = np.moveaxis(dg, 0, -1).reshape(obs.shape[0], K, 3, 3, 3)
    if "H" in want:
        ddg = (np.tensordot(CH, basis, axes=([1], [0])) * scale_c)
        ddg = np.moveaxis(ddg, 0, -1).reshape(obs.shape[0], K, 3, 3, 3, 3)
    return dg, ddg


def image_q(obs, tri, order, mu, lam, eps, want, n_quad=None):
    """Q-family at the per-observer law order (image.py:380-432)."""
    obs = np.asarray(obs, float).reshape(-1, 3)
    tri = np.asarray(tri, float).reshape(3, 3)
    nu = lam / (2.0 * (lam + mu))
    scale_c = 2.0 * (lam + mu) / (np.pi * mu * (lam + 2.0 * mu))
    fr = frame_of(tri)
    K = (order + 1) * (order + 2) // 2
    N = obs.shape[0]
    orders = (q_orders(obs, tri * REFLECT, eps, fr.L, order) if n_quad is None
              else np.full(N, int(n_quad)))
    DG = np.zeros((N, K, 3, 3, 3)) if "U" in want else None
    DDG = np.zeros((N, K, 3, 3, 3, 3)) if "H" in want else None
    for nq in np.unique(orders):
        sel = orders == nq
        dg, ddg = _q_group(obs[sel], tri, fr, order, nu, scale_c, float(eps), int(nq), want, K)
        if DG is not None:
            DG[sel] = dg
        if DDG is not None:
            DDG[sel] = ddg
    return contract(DG, DDG, unit_normal(tri), mu, lam, want)


def image_r(obs, tri, order, mu, lam, eps, want, far_field="hybrid"):
    """R-family in closed form on the reflected triangle (image.py:435-497)."""
    obs = np.asarray(obs, float).reshape(-1, 3)
    tri = np.asarray(tri, float).reshape(3, 3)
    nu = lam / (2.0 * (lam + mu))
    scale_c = 2.0 * (lam + mu) / (np.pi * mu * (lam + 2.0 * mu))
    fr_img = frame_of(tri * REFLECT)
    W, zp, _ = weighted_tables(fr_img, obs, eps, order, degrees(), far_field)
    N = obs.shape[0]
    K = next(iter(W.values())).shape[1]
    z_obs = obs[:, 2]
    cache = {}

    def moment(a, b, c, n):
        r = a + b + c
        if r == 0:
            return W[n][:, :, 0, 0]
        if (r, n) not in cache:
            cache[(r, n)] = lift(W, zp, fr_img, r, n)
        return cache[(r, n)][(slice(None), slice(None)) + (0,) * a + (1,) * b + (2,) * c]

    t = table()
    DG = DDG = None
    if "U" in want:
        DG = np.zeros((N, K, 3, 3, 3))
        for rec in t["r_dg"]:
            i, j, m, a, b, c, n = rec[:7]
            DG[:, :, i, j, m] += (_poly(rec[7], z_obs, nu) * scale_c)[:, None] * moment(a, b, c, n)
    if "H" in want:
        DDG = np.zeros((N, K, 3, 3, 3, 3))
        for rec in t["r_ddg"]:
            i, j, p, m, a, b, c, n = rec[:8]
            DDG[:, :, i, j, p, m] += (_poly(rec[8], z_obs, nu) * scale_c)[:, None] * moment(a, b, c, n)
    return contract(DG, DDG, unit_normal(tri), mu, lam, want)


def image_total(obs, tri, order, mu, lam, eps, want=("U", "H"), n_quad=None):
    """R + Q (image.py:500-514)."""
    r = image_r(obs, tri, order, mu, lam, eps, want)
    q = image_q(obs, tri, order, mu, lam, eps, want, n_quad)
    return {k: r[k] + q[k] for k in r}
