"""Derive the half-space image kernel as a monomial table (sympy, build time).

The mollified Mindlin Green's function for the half space z <= 0 is written in
Papkovich-Neuber form from Apostol's potentials (b, beta) for a unit point
force at depth z0 below the free surface (mhs follows the same route:
tools/gen_image_table.py:105-130).  Three radicals appear:

    R1 = |x - y|        (direct, full-space Kelvin part -- not tabulated here)
    R2 = |x - y'|       (y' = the source reflected in z = 0)
    Q  = R2 - (z + z0)

Each is carried as an atom with its own chain rule, so that after expansion a
term can still be classified by the radical it carries.  The image kernel is
split into

    R-family: terms carrying R2 and no Q  -> integrated in closed form,
    Q-family: terms carrying Q            -> integrated by Gauss quadrature,

and every term is rewritten in D = x - y' (so z0 = D3 - z) as

    coeff(z, nu) / (pi mu (1 - nu)) * D1^a D2^b D3^c / (R2^n Q^q).

The derivative convention matters (tools/gen_image_table.py:40-49): the image
kernel depends on z0 as well as on x - y, so d/d(src) = (-d/dx, -d/dy, +d/dz0)
and d/d(obs) = (d/dx, d/dy, d/dz).

Output: ``mhs_amd/data/mindlin_table.json`` with four record lists

    r_dg  : [i, j, m, a, b, c, n, terms]          dG_ij/dsrc_m,   R-family
    r_ddg : [i, j, p, m, a, b, c, n, terms]       d2G_ij/dobs_p dsrc_m
    q_dg  : [i, j, m, a, b, c, n, q, terms]       Q-family
    q_ddg : [i, j, p, m, a, b, c, n, q, terms]

with ``terms`` a list of [pz, pnu, num, den]: sum num/den z^pz nu^pnu.  The
record ordering is sorted by (slot, monomial), the same ordering the
reference's table uses, so the two can be compared record by record.

Run:  python -m mhs_amd.gen.mindlin_table        (takes a few minutes)
"""
from __future__ import annotations

import json
import pathlib
import sys
from collections import defaultdict

import sympy as sp

OUT = pathlib.Path(__file__).resolve().parents[1] / "data" / "mindlin_table.json"

X, Y, Z, Z0 = sp.symbols("x y z z0", real=True)
MU, NU = sp.symbols("mu nu", positive=True)
RD, RI, QQ = sp.symbols("R1 R2 Q", positive=True)
D1, D2, D3 = sp.symbols("D1 D2 D3", real=True)

# d(radical)/d(variable); the radicals never appear expanded, only as atoms.
_CHAIN = {
    RD: {X: X / RD, Y: Y / RD, Z: (Z - Z0) / RD, Z0: (Z0 - Z) / RD},
    RI: {X: X / RI, Y: Y / RI, Z: (Z + Z0) / RI, Z0: (Z + Z0) / RI},
    QQ: {X: X / RI, Y: Y / RI, Z: (Z + Z0) / RI - 1, Z0: (Z + Z0) / RI - 1},
}
_SCALE = 1 / (sp.pi * MU * (1 - NU))


def diff(expr, var):
    """Total derivative with R1, R2, Q treated as functions of (x, y, z, z0)."""
    total = sp.diff(expr, var)
    for atom, rule in _CHAIN.items():
        total += sp.diff(expr, atom) * rule[var]
    return total


def d_obs(expr, p):
    return diff(expr, (X, Y, Z)[p])


def d_src(expr, m):
    if m == 2:
        return diff(expr, Z0)
    return -diff(expr, (X, Y)[m])


def apostol(force: int):
    """Papkovich-Neuber potentials (b, beta) of a unit force along axis
    ``force`` at depth z0 (Apostol's half-space form)."""
    k = 1 / (4 * sp.pi * MU)
    a0 = -Z0
    if force == 2:
        b = [sp.S(0), sp.S(0),
             k * (1 / RD + (3 - 4 * NU) / RI + 2 * Z0 * (Z + Z0) / RI ** 3)]
        beta = k * (a0 / RD + (3 - 4 * NU) * a0 / RI
                    - 4 * (1 - NU) * (1 - 2 * NU) * sp.log(QQ))
        return b, beta
    t = (X, Y)[force]
    b = [sp.S(0), sp.S(0), sp.S(0)]
    b[force] = k * (1 / RD + 1 / RI)
    b[2] = 2 * k * (a0 / RI ** 2 - (1 - 2 * NU) / QQ) * t / RI
    beta = 2 * k * (1 - 2 * NU) * (a0 / RI - (1 - 2 * NU)) * t / QQ
    return b, beta


def greens():
    """G[i][j]: displacement i from a unit force j."""
    G = [[None] * 3 for _ in range(3)]
    for j in range(3):
        b, beta = apostol(j)
        phi = X * b[0] + Y * b[1] + Z * b[2] + beta
        for i in range(3):
            G[i][j] = b[i] - d_obs(phi, i) / (4 * (1 - NU))
    return G


def split(expr):
    """(R-family, Q-family) parts of an expression, by the radical carried."""
    r = sp.S(0)
    q = sp.S(0)
    for term in sp.Add.make_args(sp.expand(expr)):
        if term.has(QQ):
            q += term
        elif term.has(RI):
            r += term
    return r, q


def monomials(expr, with_q):
    """{(a, b, c, n[, q]): coeff} in D = x - y'; z0 -> D3 - z."""
    e = sp.expand(expr.subs(Z0, D3 - Z).subs({X: D1, Y: D2}))
    acc = defaultdict(lambda: sp.S(0))
    for term in sp.Add.make_args(e):
        num, den = sp.fraction(sp.cancel(sp.together(term)))
        deg = {s: sp.degree(num, s) - sp.degree(den, s)
               for s in (D1, D2, D3, RI, QQ)}
        a, b, c = int(deg[D1]), int(deg[D2]), int(deg[D3])
        n, q = int(-deg[RI]), int(-deg[QQ])
        assert min(a, b, c) >= 0, term
        if not with_q:
            assert q == 0 and n % 2 == 1, term
        coeff = sp.cancel(term * RI ** n * QQ ** q / (D1 ** a * D2 ** b * D3 ** c))
        assert not (coeff.free_symbols & {D1, D2, D3, RI, QQ}), coeff
        acc[(a, b, c, n, q) if with_q else (a, b, c, n)] += coeff
    out = {}
    for key, val in acc.items():
        val = sp.cancel(val)
        if val != 0:
            out[key] = val
    return out


def poly_terms(coeff):
    """coeff / (pi mu (1-nu)) as [[pz, pnu, num, den], ...], sorted."""
    rest = sp.cancel(coeff / _SCALE)
    assert rest.free_symbols <= {Z, NU}, rest
    poly = sp.Poly(sp.expand(rest), Z, NU)
    terms = []
    for (pz, pnu), c in sorted(poly.terms()):
        c = sp.Rational(c)
        terms.append([int(pz), int(pnu), int(c.p), int(c.q)])
    return terms


def build():
    G = greens()
    r_dg, r_ddg, q_dg, q_ddg = [], [], [], []
    for i in range(3):
        for j in range(3):
            for m in range(3):
                dr, dq = split(d_src(G[i][j], m))
                for key, co in sorted(monomials(dr, False).items()):
                    r_dg.append([i, j, m, *key, poly_terms(co)])
                for key, co in sorted(monomials(dq, True).items()):
                    q_dg.append([i, j, m, *key, poly_terms(co)])
                for p in range(3):
                    er, _ = split(d_obs(dr, p))
                    _, eq = split(d_obs(dq, p))
                    for key, co in sorted(monomials(er, False).items()):
                        r_ddg.append([i, j, p, m, *key, poly_terms(co)])
                    for key, co in sorted(monomials(eq, True).items()):
                        q_ddg.append([i, j, p, m, *key, poly_terms(co)])
            print(f"  G[{i}][{j}] done", flush=True)
    return {"r_dg": r_dg, "r_ddg": r_ddg, "q_dg": q_dg, "q_ddg": q_ddg}


def main() -> int:
    table = build()
    for k, v in table.items():
        print(f"{k}: {len(v)} records")
    OUT.parent.mkdir(parents=True, exist_ok=True)
    OUT.write_text(json.dumps(table, separators=(",", ":")) + "\n")
    print(f"wrote {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
