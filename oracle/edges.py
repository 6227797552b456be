"""Edge integrals P_k^m = int u^k (u^2 + rho^2)^(-m/2) du and the solid angle.

Restates src/mhs/fullspace/primitives.py:66-528: four cancellation-free
regimes chosen per edge (closed form, small-|u| series, large-|u| series,
rho = 0), 32 series terms, thresholds SERIES_U_OVER_RHO = 2 and
SMALL_U_OVER_RHO = 0.5 (fullspace/defaults.py:26-33).
"""
from __future__ import annotations

import functools
from fractions import Fraction

import numpy as np

SERIES_U_OVER_RHO = 2.0
SMALL_U_OVER_RHO = 0.5
SERIES_TERMS = 32


@functools.lru_cache(maxsize=None)
def binom_half(m: int, j: int) -> float:
    """C(-m/2, j), exact then rounded (primitives.py:65-84)."""
    a = Fraction(-m, 2)
    out = Fraction(1)
    for i in range(j):
        out *= (a - i) / (i + 1)
    return float(out)


class _Closed:
    """Closed-form differences for one edge batch (primitives.py:131-212)."""

    def __init__(self, ua, ub, rho2):
        self.ua, self.ub, self.rho2 = ua, ub, rho2
        self.Ra = np.sqrt(ua * ua + rho2)
        self.Rb = np.sqrt(ub * ub + rho2)
        self.dR = (ub - ua) * (ub + ua) / (self.Ra + self.Rb)
        self.J, self.K, self.P = {}, {}, {}

    def dJ(self, m):
        if m in self.J:
            return self.J[m]
        ua, ub, Ra, Rb, rho2, dR = self.ua, self.ub, self.Ra, self.Rb, self.rho2, self.dR
        with np.errstate(divide="ignore", invalid="ignore"):
            if m == 1:
                pos = ua >= 0.0
                neg = ub <= 0.0
                mix = ~(pos | neg)
                out = np.empty_like(ua)
                out[pos] = np.log1p(((ub - ua) + dR)[pos] / (ua + Ra)[pos])
                out[neg] = np.log1p(((ub - ua) - dR)[neg] / (Rb - ub)[neg])
                out[mix] = np.log(((ub + Rb) * (Ra - ua))[mix] / rho2[mix])
            elif m == 3:
                same = (ua * ub) > 0.0
                out = np.empty_like(ua)
                den = Ra * Rb * (ub * Ra + ua * Rb)
                out[same] = ((ub - ua) * (ub + ua))[same] / den[same]
                out[~same] = (ub / Rb - ua / Ra)[~same] / rho2[~same]
            elif m >= 5:
                mm = m - 2
                out = (ub * Rb ** (-mm) - ua * Ra ** (-mm) + (mm - 1) * self.dJ(mm)) / (mm * rho2)
            elif m == -1:
                duR = 0.5 * (ub - ua) * (Ra + Rb) + 0.5 * (ua + ub) * dR
                out = 0.5 * (duR + rho2 * self.dJ(1))
            else:
                p = -m
                out = (ub * Rb ** p - ua * Ra ** p) / (p + 1) + p * rho2 / (p + 1) * self.dJ(m + 2)
        self.J[m] = out
        return out

    def dK(self, m):
        if m in self.K:
            return self.K[m]
        Ra, Rb, dR = self.Ra, self.Rb, self.dR
        q = 2 - m
        if q > 0:
            S = sum(Ra ** i * Rb ** (q - 1 - i) for i in range(q))
            out = dR * S / q
        else:
            p = -q
            S = sum(Ra ** i * Rb ** (p - 1 - i) for i in range(p))
            out = dR * S / (p * (Ra * Rb) ** p)
        self.K[m] = out
        return out

    def dP(self, k, m):
        key = (k, m)
        if key not in self.P:
            if k == 0:
                self.P[key] = self.dJ(m)
            elif k == 1:
                self.P[key] = self.dK(m)
            else:
                self.P[key] = self.dP(k - 2, m - 2) - self.rho2 * self.dP(k - 2, m)
        return self.P[key]


def _dpow(xa, xb, q, log_r, close):
    """xb^q - xa^q, expm1 form when the endpoints are close
    (primitives.py:215-224)."""
    if q == 0:
        return log_r
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        return np.where(close, xa ** q * np.expm1(q * log_r), xb ** q - xa ** q)


def _large_u(k, m, ua, ub, rho2):
    """Large-|u| series, 0 < ua < ub (primitives.py:227-256 / 293-348)."""
    x = rho2 / (ua * ua)
    r = ub / ua
    log_r = np.log1p((ub - ua) / ua)
    close = r < 2.0
    ones = np.ones_like(ua)
    s = k - m + 1
    out = np.zeros_like(ua)
    xj = np.ones_like(ua)
    for j in range(SERIES_TERMS):
        q = s - 2 * j
        t = _dpow(ones, r, q, log_r, close)
        if q != 0:
            t = t / q
        out = out + binom_half(m, j) * xj * t
        xj = xj * x
    return ua ** s * out


def _small_u(k, m, ua, ub, rho2):
    """Small-|u| series, max|u| <= rho/2 (primitives.py:259-290 / 351-398)."""
    rho = np.sqrt(rho2)
    ya, yb = ua / rho, ub / rho
    same = (ua * ub) > 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        log_r = np.where(same, np.log1p((ub - ua) / ua), 0.0)
        ratio = np.where(same, ub / ua, np.inf)
    close = same & (np.abs(ratio) < 2.0) & (np.abs(ratio) > 0.5)
    out = np.zeros_like(ua)
    for j in range(SERIES_TERMS):
        pw = k + 2 * j + 1
        with np.errstate(over="ignore", invalid="ignore"):
            d = np.where(close, ya ** pw * np.expm1(pw * log_r), yb ** pw - ya ** pw)
        out = out + binom_half(m, j) * d / pw
    return rho ** (k + 1 - m) * out


def edge_table(ua, ub, rho2, spec):
    """{m: (N, kmax+1)} edge integrals for ub > ua (primitives.py:431-504).
    rho2 > 0 always here: eps > 0 is enforced at the boundary, so the rho = 0
    regime of the reference (primitives.py:401-428) is unreachable and is
    refused rather than restated."""
    ua = np.asarray(ua, float)
    ub = np.asarray(ub, float)
    rho2 = np.asarray(rho2, float)
    if np.any(rho2 <= 0.0):
        raise ValueError("rho = 0 edge: only reachable with eps = 0")
    rho = np.sqrt(rho2)
    out = {m: np.empty((ua.shape[0], kmax + 1)) for m, kmax in spec.items()}
    umax = np.maximum(np.abs(ua), np.abs(ub))
    umin = np.minimum(np.abs(ua), np.abs(ub))
    small = umax <= SMALL_U_OVER_RHO * rho
    ser = ((ua * ub) > 0.0) & (umin >= SERIES_U_OVER_RHO * rho)
    closed = ~small & ~ser
    if np.any(closed):
        cf = _Closed(ua[closed], ub[closed], rho2[closed])
        for m, kmax in spec.items():
            for k in range(kmax + 1):
                out[m][closed, k] = cf.dP(k, m)
    if np.any(small):
        for m, kmax in spec.items():
            for k in range(kmax + 1):
                out[m][small, k] = _small_u(k, m, ua[small], ub[small], rho2[small])
    if np.any(ser):
        a, b = ua[ser], ub[ser]
        neg = a < 0.0
        ap = np.where(neg, -b, a)
        bp = np.where(neg, -a, b)
        for m, kmax in spec.items():
            for k in range(kmax + 1):
                v = _large_u(k, m, ap, bp, rho2[ser])
                if k % 2 == 1:
                    v = np.where(neg, -v, v)
                out[m][ser, k] = v
    return out


def solid_angle(v1, v2, v3, obs):
    """Van Oosterom-Strackee signed solid angle (primitives.py:511-528)."""
    obs = np.asarray(obs, float).reshape(-1, 3)
    r1, r2, r3 = v1[None] - obs, v2[None] - obs, v3[None] - obs
    R1 = np.linalg.norm(r1, axis=1)
    R2 = np.linalg.norm(r2, axis=1)
    R3 = np.linalg.norm(r3, axis=1)
    num = np.einsum("ni,ni->n", r1, np.cross(r2, r3))
    den = (R1 * R2 * R3 + R3 * np.einsum("ni,ni->n", r1, r2)
           + R1 * np.einsum("ni,ni->n", r2, r3) + R2 * np.einsum("ni,ni->n", r1, r3))
    return 2.0 * np.arctan2(num, den)
