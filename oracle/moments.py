"""Per-node weighted moment tables W_k[n][a, b] = int_T N_k xi1^a xi2^b R^-n dS.

Restates src/mhs/fullspace/moments.py: degree bookkeeping (63-116), the
closed-form MomentTable (171-363), weighted_from_table (370-398), the far-field
Gauss producer (427-498) and the hybrid switch of weighted_tables (515-606):
closed form where sqrt(|x - c|^2 + eps^2) <= D_STAR L, Gauss 12^2 up to
D_STAR_DISTANT L and 8^2 beyond (fullspace/defaults.py:20-23).

eps > 0 always (the boundary refuses eps <= 0), so h_eps > 0 and the
reference's on-plane floor machinery (moments.py:119-164) never engages.
"""
from __future__ import annotations

from math import comb

import numpy as np

from .edges import edge_table, solid_angle
from .geometry import Frame, gauss_triangle, nodal_at, shape_coefficients

D_STAR = 10.0
D_STAR_DISTANT = 40.0
FAR_GAUSS_N = 12
FAR_GAUSS_N_DISTANT = 8


def kernel_degrees(order: int, want) -> dict:
    """moments.py:63-91 (slip kernels only)."""
    need = {}

    def add(n, d):
        need[n] = max(need.get(n, -1), d)
    if "U" in want:
        add(3, 1 + order); add(5, 3 + order)
    if "H" in want:
        add(3, order); add(5, 2 + order); add(7, 4 + order)
    if "E" in want:
        add(7, order)
    return need


def close_degrees(need: dict) -> dict:
    """Lower-n degrees the recursion consumes (moments.py:94-116)."""
    deg = dict(need)
    n = max(deg)
    while True:
        d = deg.get(n, -1)
        if d >= 2:
            deg[n - 2] = max(deg.get(n - 2, -1), d - 2)
        if n - 2 < min(deg):
            break
        n -= 2
    deg[3] = max(deg.get(3, -1), 0)
    for n in list(deg):
        if n >= 5:
            for m in range(3, n, 2):
                deg[m] = max(deg.get(m, -1), 0)
        if n <= 1:
            for m in range(n, 4, 2):
                deg[m] = max(deg.get(m, -1), 0)
    return {n: d for n, d in deg.items() if d >= 0}


def edge_spec(deg: dict) -> dict:
    """{m: kmax} of edge integrals the table needs (moments.py:221-233)."""
    spec = {}
    for n, d in deg.items():
        if d >= 1:
            spec[n - 2] = max(spec.get(n - 2, -1), d - 1)
    for n in deg:
        if n >= 5:
            spec[n - 2] = max(spec.get(n - 2, -1), 0)
        if n <= 1:
            spec[n] = max(spec.get(n, -1), 0)
    return spec


def raw_moments(fr: Frame, obs, eps: float, degrees: dict):
    """(z, X, M) with M[n] (N, D+1, D+1) the raw moments M_n^(a,b)
    (moments.py:181-355)."""
    obs = np.asarray(obs, float).reshape(-1, 3)
    N = obs.shape[0]
    z, X = fr.to_plane(obs)
    h2 = z * z + eps ** 2
    h = np.sqrt(h2)
    deg = close_degrees(degrees)
    p = fr.p[None, :, :] - X[:, None, :]
    obs_eff = fr.centroid[None] + X[:, :1] * fr.e1 + X[:, 1:] * fr.e2 + h[:, None] * fr.nhat[None]
    I3 = -solid_angle(fr.v[0], fr.v[1], fr.v[2], obs_eff) / h
    spec = edge_spec(deg)
    BN1, BN2, BD = {}, {}, {}
    for e in range(3):
        pa = p[:, e, :]
        pb = p[:, (e + 1) % 3, :]
        ev = pb - pa
        t = ev / np.linalg.norm(ev, axis=1)[:, None]
        c1, c2 = t[:, 1], -t[:, 0]                     # outward normal
        s1, s2 = t[:, 0], t[:, 1]
        dp = pa[:, 0] * c1 + pa[:, 1] * c2
        ua = pa[:, 0] * s1 + pa[:, 1] * s2
        ub = pb[:, 0] * s1 + pb[:, 1] * s2
        tab = edge_table(ua, ub, dp * dp + h2, spec)
        nz = dp != 0.0

        def term(coeff, col, pw):
            # moments.py:263-280: a zero dp kills the term exactly
            if pw == 0:
                return coeff * col
            out = np.zeros(N)
            out[nz] = coeff[nz] * col[nz]
            return out
        for m, kmax in spec.items():
            P = tab[m]
            BD[m] = BD.get(m, 0.0) + term(dp, P[:, 0], 1)
            for a in range(kmax + 1):
                for b in range(kmax + 1 - a):
                    val = np.zeros(N)
                    for i in range(a + 1):
                        for j in range(b + 1):
                            pw = a + b - i - j
                            coeff = (comb(a, i) * comb(b, j) * dp ** pw
                                     * c1 ** (a - i) * s1 ** i * c2 ** (b - j) * s2 ** j)
                            val += term(coeff, P[:, i + j], pw)
                    BN1[(a, b, m)] = BN1.get((a, b, m), 0.0) + c1 * val
                    BN2[(a, b, m)] = BN2.get((a, b, m), 0.0) + c2 * val
    M = {n: np.zeros((N, d + 1, d + 1)) for n, d in deg.items()}
    M[3][:, 0, 0] = I3
    n = 3
    while n + 2 in M:                                   # upward seeds
        M[n + 2][:, 0, 0] = (BD[n] + (n - 2) * M[n][:, 0, 0]) / (n * h2)
        n += 2
    n = 1
    while n in M:                                       # downward seeds
        M[n][:, 0, 0] = (BD[n] - n * h2 * M[n + 2][:, 0, 0]) / (2.0 - n)
        n -= 2
    for n in sorted(M):                                 # horizontal recursion
        d = deg[n]
        m = n - 2
        for q in range(1, d + 1):
            for a in range(q, -1, -1):
                b = q - a
                if a >= 1:
                    val = -BN1[(a - 1, b, m)]
                    if a >= 2:
                        val = val + (a - 1) * M[m][:, a - 2, b]
                else:
                    val = -BN2[(0, b - 1, m)]
                    if b >= 2:
                        val = val + (b - 1) * M[m][:, 0, b - 2]
                M[n][:, a, b] = val / (n - 2)
    return z, X, M


def weighted_near(fr: Frame, obs, eps, order, need0):
    """Closed-form weighted tables (moments.py:370-398 over 181-355)."""
    need = {n: d + order for n, d in need0.items()}
    z, X, M = raw_moments(fr, obs, eps, need)
    c = shape_coefficients(fr, order, X)
    N, K, D1, _ = c.shape
    W = {}
    for n, d in need0.items():
        Wn = np.zeros((N, K, d + 1, d + 1))
        for a in range(d + 1):
            for b in range(d + 1 - a):
                acc = np.zeros((N, K))
                for ap in range(D1):
                    for bp in range(D1 - ap):
                        cc = c[:, :, ap, bp]
                        if not np.any(cc):
                            continue
                        acc += cc * M[n][:, a + ap, b + bp][:, None]
                Wn[:, :, a, b] = acc
        W[n] = Wn
    return W, z, X


def weighted_far(fr: Frame, obs, eps, order, need0, ng):
    """Gauss ng^2 weighted tables (moments.py:427-498)."""
    obs = np.asarray(obs, float).reshape(-1, 3)
    z, X = fr.to_plane(obs)
    h2 = z * z + eps ** 2
    x1, x2, w = gauss_triangle(ng)
    lam = np.stack([1.0 - x1 - x2, x1, x2], axis=1)
    eta = lam @ fr.p
    wq = w * (2.0 * fr.area)
    Nq = nodal_at(fr, order, eta)
    xi0 = eta[None, :, 0] - X[:, None, 0]
    xi1 = eta[None, :, 1] - X[:, None, 1]
    R2 = xi0 * xi0 + xi1 * xi1 + h2[:, None]
    inv = 1.0 / np.sqrt(R2)
    inv2 = inv * inv
    rpow = {}
    cur = inv
    for n in range(1, max(need0) + 1, 2):
        rpow[n] = cur
        cur = cur * inv2
    dmax = max(need0.values())
    p0, p1 = [np.ones_like(xi0)], [np.ones_like(xi1)]
    for _ in range(dmax):
        p0.append(p0[-1] * xi0)
        p1.append(p1[-1] * xi1)
    W = {}
    for n, d in need0.items():
        Wn = np.zeros((X.shape[0], Nq.shape[1], d + 1, d + 1))
        wr = wq[None, :] * rpow[n]
        for a in range(d + 1):
            wa = wr * p0[a]
            for b in range(d + 1 - a):
                Wn[:, :, a, b] = (wa * p1[b]) @ Nq
        W[n] = Wn
    return W, z, X


def weighted_tables(fr: Frame, obs, eps, order, need0, far_field="hybrid"):
    """Hybrid switch (moments.py:515-606): returns (W, z, X)."""
    obs = np.asarray(obs, float).reshape(-1, 3)
    N = obs.shape[0]
    K = (order + 1) * (order + 2) // 2
    D = np.sqrt(np.sum((obs - fr.centroid) ** 2, axis=1) + eps ** 2)
    if far_field == "analytic":
        near = np.ones(N, bool)
    elif far_field == "quadrature":
        near = np.zeros(N, bool)
    else:
        near = D <= D_STAR * fr.L
    W = {n: np.empty((N, K, d + 1, d + 1)) for n, d in need0.items()}
    z = np.empty(N)
    X = np.empty((N, 2))
    if np.any(near):
        Wn, zn, Xn = weighted_near(fr, obs[near], eps, order, need0)
        for n in W:
            W[n][near] = Wn[n]
        z[near], X[near] = zn, Xn
    far = ~near
    if np.any(far):
        ng = np.where(D > D_STAR_DISTANT * fr.L, FAR_GAUSS_N_DISTANT, FAR_GAUSS_N)
        for g in (FAR_GAUSS_N, FAR_GAUSS_N_DISTANT):
            sel = far & (ng == g)
            if np.any(sel):
                Wf, zf, Xf = weighted_far(fr, obs[sel], eps, order, need0, g)
                for n in W:
                    W[n][sel] = Wf[n]
                z[sel], X[sel] = zf, Xf
    return W, z, X
