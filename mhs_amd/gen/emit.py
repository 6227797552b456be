"""Emit the generated HIP device code for the image families.

Reads ``mhs_amd/data/mindlin_table.json`` (derived by ``mindlin_table.py``)
and writes ``mhs_amd/csrc/gen_image.inc`` containing

* ``MHS_NCOEF`` and the host routine ``mhs_fill_coeffs(mu, lam, c)`` that
  turns the exact rational records into per-material FP64 coefficients
  (nu and the common scale 2 (lam + mu) / (pi mu (lam + 2 mu)) folded in,
  src/mhs/kernels/image.py:95-104, 276-297, 405-406);
* ``r_ddg`` / ``r_dg``: the R-family, straight-line code that lifts the
  image-frame weighted moment tables W[n][u][v] to the Cartesian monomial
  moments int N D1^a D2^b D3^c / R2^n (D = -xi1 e1' - xi2 e2' + z' n') and
  accumulates the records into DDG (81) / DG (27);
* ``q_blocked`` / ``r_blocked``: the contracted image series, register-blocked
  over quadrature points; ``qmom_*``: the Q-family's moments-first form.

The traversal is depth-first over the monomial tree, so only one root-to-leaf
path of expansion polynomials is live at a time.

Run:  python -m mhs_amd.gen.emit
"""
from __future__ import annotations

import json
import pathlib
from collections import defaultdict
from fractions import Fraction

ROOT = pathlib.Path(__file__).resolve().parents[1]
TABLE = ROOT / "data" / "mindlin_table.json"
OUT = ROOT / "csrc" / "gen_image.inc"


def tidx(u, v):
    """Triangular index of the monomial xi1^u xi2^v."""
    s = u + v
    return s * (s + 1) // 2 + v


def slot_of(idx):
    s = 0
    for i in idx:
        s = s * 3 + i
    return s


class Emitter:
    def __init__(self):
        self.lines = []
        self.n_coef = 0
        self.coef_src = []          # (kind, list of (pnu, Fraction)) per coefficient
        self.rec_ks = {}            # table name -> per record {pz: coefficient index}

    def coef(self, terms):
        """Register one coefficient sum_t num/den nu^pnu; return its index."""
        self.coef_src.append(terms)
        self.n_coef += 1
        return self.n_coef - 1

    def w(self, s=""):
        self.lines.append(s)


def emit_r(em: Emitter, name, records, n_ix, n_slots):
    """R-family: DDG (n_ix = 4) or DG (n_ix = 3)."""
    # group records by monomial (a, b, c, n)
    by_mono = defaultdict(list)
    for rec in records:
        slot = slot_of(rec[:n_ix])
        a, b, c, n = rec[n_ix:n_ix + 4]
        by_pz = defaultdict(list)
        for pz, pnu, num, den in rec[n_ix + 4]:
            by_pz[pz].append((pnu, Fraction(num, den)))
        ks = {pz: em.coef(t) for pz, t in sorted(by_pz.items())}
        em.rec_ks.setdefault(name, []).append(ks)
        by_mono[(a, b, c)].append((n, slot, ks))
    # monomial tree: parent drops one factor (z first, then y, then x)
    need = set(by_mono)
    for m in list(need):
        a, b, c = m
        while a + b + c > 0:
            if c > 0:
                c -= 1
            elif b > 0:
                b -= 1
            else:
                a -= 1
            need.add((a, b, c))

    def parent(m):
        a, b, c = m
        if c > 0:
            return (a, b, c - 1), 2
        if b > 0:
            return (a, b - 1, c), 1
        return (a - 1, b, c), 0
    children = defaultdict(list)
    for m in need:
        if m != (0, 0, 0):
            children[parent(m)[0]].append(m)
    em.w(f"__host__ __device__ __forceinline__ void {name}(const double* __restrict__ W3,")
    em.w("    const double* __restrict__ W5, const double* __restrict__ W7,")
    em.w("    const double* __restrict__ W9, const double* al, const double* be,")
    em.w("    const double* ga, double z, const double* __restrict__ C, double* acc)")
    em.w("{")
    em.w("  const double z2 = z * z;")
    Wn = {3: "W3", 5: "W5", 7: "W7", 9: "W9"}

    def poly_name(m, u, v):
        return f"p{m[0]}{m[1]}{m[2]}_{u}_{v}"

    def visit(m, depth):
        r = sum(m)
        ind = "  " * (depth + 1)
        if m == (0, 0, 0):
            pass
        else:
            par, axis = parent(m)
            for s in range(r + 1):
                for v in range(s + 1):
                    u = s - v
                    terms = []
                    if u > 0 and (u - 1) + v <= r - 1:
                        terms.append(f"al[{axis}] * {poly_name(par, u - 1, v)}" if par != (0, 0, 0)
                                     else (f"al[{axis}]" if (u - 1, v) == (0, 0) else None))
                    if v > 0 and u + (v - 1) <= r - 1:
                        terms.append(f"be[{axis}] * {poly_name(par, u, v - 1)}" if par != (0, 0, 0)
                                     else (f"be[{axis}]" if (u, v - 1) == (0, 0) else None))
                    if u + v <= r - 1:
                        terms.append(f"ga[{axis}] * {poly_name(par, u, v)}" if par != (0, 0, 0)
                                     else (f"ga[{axis}]" if (u, v) == (0, 0) else None))
                    terms = [t for t in terms if t is not None]
                    em.w(f"{ind}const double {poly_name(m, u, v)} = {' + '.join(terms) if terms else '0.0'};")
        by_n = defaultdict(list)
        for n, slot, ks in by_mono.get(m, []):
            by_n[n].append((slot, ks))
        for n in sorted(by_n):
            if m == (0, 0, 0):
                expr = f"{Wn[n]}[0]"
            else:
                expr = " + ".join(f"{poly_name(m, s - v, v)} * {Wn[n]}[{tidx(s - v, v)}]"
                                  for s in range(r + 1) for v in range(s + 1))
            em.w(f"{ind}{{ const double mm = {expr};")
            em.w(f"{ind}  const double mz = z * mm, mzz = z2 * mm;")
            for slot, ks in by_n[n]:
                parts = []
                for pz, k in ks.items():
                    parts.append(f"C[{k}] * {('mm', 'mz', 'mzz')[pz]}")
                em.w(f"{ind}  acc[{slot}] += {' + '.join(parts)};")
            em.w(f"{ind}}}")
        kids = sorted(children.get(m, []))
        for ch in kids:
            em.w(f"{ind}{{")
            visit(ch, depth + 1)
            em.w(f"{ind}}}")
    visit((0, 0, 0), 0)
    em.w("}")
    em.w()


def emit_q(em: Emitter, name, records, n_ix):
    """Register the Q-family records' coefficients (one per record, in record
    order); the kernels evaluate the family through the contracted and the
    moments-first forms emitted below."""
    for rec in records:
        terms = [(pnu, Fraction(num, den)) for pz, pnu, num, den in rec[n_ix + 5]]
        assert all(pz == 0 for pz, *_ in rec[n_ix + 5])
        em.coef(terms)


SYM = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (0, 2): 4, (2, 0): 4, (1, 2): 5, (2, 1): 5}


def emit_blocked_sum(em: Emitter, name, monos, groups, combine):
    """Register-blocked sum_k kap[o * M + k] * d1^a d2^b d3^c * radial_g for
    NPT points: the spatial monomials by an x-major product chain (one
    multiply each), each feeding its radial group's running sum acc[g];
    ``combine`` emits the per-point radial combination into S.  ``monos`` is
    [(a, b, c, radial key)] in kappa order (spatial part ascending)."""
    M = len(monos)
    gidx = {g: i for i, g in enumerate(groups)}
    dparts = sorted({m[:3] for m in monos})
    by_d = defaultdict(list)
    for k, m in enumerate(monos):
        by_d[m[:3]].append((k, m))
    # the body is emitted twice: reading kappa from memory (a pointer), and
    # from any indexable KapT (``{name}_k``: kappa formed on the fly)
    out_em = em
    em = Emitter()
    em.w(f"  T acc[NOUT][NPT][{len(groups)}];")
    em.w("  T xa[NPT], xy[NPT], d[NPT];")
    init = set()
    amax = max(a for a, _, _ in dparts)
    for a in range(amax + 1):
        da = [p for p in dparts if p[0] == a]
        if not da:
            continue
        if a == 1:
            em.w("#pragma unroll")
            em.w("  for (int j = 0; j < NPT; ++j) xa[j] = x[j];")
        elif a > 1:
            em.w("#pragma unroll")
            em.w("  for (int j = 0; j < NPT; ++j) xa[j] *= x[j];")
        bmax = max(p[1] for p in da)
        for b in range(bmax + 1):
            db = [p for p in da if p[1] == b]
            xa_one = a == 0
            if b == 0:
                xy_expr = None if xa_one else "xa[j]"
            elif b == 1:
                xy_expr = "y[j]" if xa_one else "xa[j] * y[j]"
            else:
                xy_expr = "xy[j] * y[j]"
            xy_one = xy_expr is None
            if xy_expr is not None:
                em.w("#pragma unroll")
                em.w(f"  for (int j = 0; j < NPT; ++j) xy[j] = {xy_expr};")
            if not db:
                continue
            cmax = max(p[2] for p in db)
            for c in range(cmax + 1):
                if c == 0:
                    one = xy_one
                    src = None if one else "xy"
                else:
                    one = False
                    expr = ("z[j]" if xy_one else "xy[j] * z[j]") if c == 1 else "d[j] * z[j]"
                    em.w("#pragma unroll")
                    em.w(f"  for (int j = 0; j < NPT; ++j) d[j] = {expr};")
                    src = "d"
                ms = by_d.get((a, b, c), [])
                if not ms:
                    continue
                em.w("#pragma unroll")
                em.w("  for (int o = 0; o < NOUT; ++o) {")
                for k, _ in ms:
                    em.w(f"    const T k{k} = kap[o * {M} + {k}];")
                em.w("#pragma unroll")
                em.w("    for (int j = 0; j < NPT; ++j) {")
                for k, m in ms:
                    g = gidx[m[3]]
                    tgt = f"acc[o][j][{g}]"
                    if g not in init:
                        em.w(f"      {tgt} = {f'k{k}' if one else f'k{k} * {src}[j]'};")
                    else:
                        em.w(f"      {tgt} = {f'{tgt} + k{k}' if one else f'fma(k{k}, {src}[j], {tgt})'};")
                em.w("    }")
                em.w("  }")
                for _, m in ms:
                    init.add(gidx[m[3]])
    assert len(init) == len(groups)
    em.w("#pragma unroll")
    em.w("  for (int j = 0; j < NPT; ++j) {")
    combine(em, gidx)
    em.w("  }")
    body = em.lines
    em = out_em
    em.w("template <typename T, int NPT, int NOUT>")
    em.w(f"__device__ __forceinline__ void {name}(const T* __restrict__ kap, const T* x, const T* y,")
    em.w("    const T* z, const T* ir, const T* iq, T* s)")
    em.w("{")
    em.lines += body
    em.w("}")
    em.w()
    em.w("template <typename T, int NPT, int NOUT, typename KapT>")
    em.w(f"__device__ __forceinline__ void {name}_k(const KapT& kap, const T* x, const T* y,")
    em.w("    const T* z, const T* ir, const T* iq, T* s)")
    em.w("{")
    em.lines += body
    em.w("}")
    em.w()


def _walk_order(monos):
    """kappa order of the blocked sum: spatial part (x-major), then radial."""
    return sorted(monos, key=lambda m: (m[:3], m[3]))


def _csr(em, prefix, keyed, n_keys, cols):
    """CSR over keys k * 6 + v of entries (key, *cols) (sorted by key)."""
    keyed.sort()
    ptr = [0] * (n_keys + 1)
    for e in keyed:
        ptr[e[0] + 1] += 1
    for t in range(n_keys):
        ptr[t + 1] += ptr[t]
    em.w(f"__constant__ short {prefix}_ptr[{len(ptr)}] = {{{', '.join(map(str, ptr))}}};")
    for ci, (cname, ctype) in enumerate(cols):
        vals = ", ".join(str(e[ci + 1]) for e in keyed)
        em.w(f"__constant__ {ctype} {prefix}_{cname}[{len(keyed)}] = {{{vals}}};")


def emit_q_kappa(em: Emitter, records, n_ix):
    """Contracted Q-family: per pair, kappa_k = sum over the records of
    monomial k of C[rec] * G[slot] (G = the pair's output functional on the
    81 DDG slots, G[i, j, p, m] = E[i][p] F[j][m] with E the receiver and F
    the source half, both symmetric); factored as kappa_k = sum_v F_v B[k][v]
    with the receiver table B[k][v] = sum over the records of k whose (j, m)
    is the symmetric index v of C[rec] E[i][p].  Per quadrature point
    sum_k kappa_k d1^a d2^b d3^c R^-n Q^-q, register-blocked; the radial
    groups (n, q) combine by nested Horner in 1/Q then 1/R."""
    by_mono = defaultdict(list)
    for r, rec in enumerate(records):
        by_mono[tuple(rec[n_ix:n_ix + 3]) + (tuple(rec[n_ix + 3:n_ix + 5]),)].append((slot_of(rec[:n_ix]), r))
    monos = _walk_order(list(by_mono))
    groups = sorted({m[3] for m in monos})
    kidx = {m: i for i, m in enumerate(monos)}
    base = em.q_ddg_base
    keyed = []
    for m in monos:
        for slot, r in by_mono[m]:
            i, j, p, mm = slot // 27, (slot // 9) % 3, (slot // 3) % 3, slot % 3
            keyed.append((kidx[m] * 6 + SYM[(j, mm)], base + r, i * 3 + p))
    em.w(f"#define MHS_QM {len(monos)}")
    em.w(f"#define MHS_QREC {len(keyed)}")
    em.w("// Q B-table CSR over (monomial k, symmetric source index v): record")
    em.w("// coefficient index and receiver index i * 3 + p.")
    _csr(em, "mhs_qb", keyed, len(monos) * 6, [("coef", "short"), ("ip", "unsigned char")])
    ns = sorted({g[0] for g in groups})
    assert ns == list(range(ns[0], ns[-1] + 1)) and ns[0] == 0
    qmins = {n: min(q for nn, q in groups if nn == n) for n in ns}

    def combine(em, gidx):
        em.w("    const T q1 = iq[j];")
        for k in range(2, max(qmins.values()) + 1):
            em.w(f"    const T q{k} = q{k - 1} * iq[j];")
        em.w("#pragma unroll")
        em.w("    for (int o = 0; o < NOUT; ++o) {")
        em.w("      T S, t;")
        for n in reversed(ns):
            qs = sorted(q for nn, q in groups if nn == n)
            assert qs == list(range(qs[0], qs[-1] + 1))
            em.w(f"      t = acc[o][j][{gidx[(n, qs[-1])]}];")
            for q in reversed(qs[:-1]):
                em.w(f"      t = fma(t, q1, acc[o][j][{gidx[(n, q)]}]);")
            if n == ns[-1]:
                em.w(f"      S = t * q{qmins[n]};")
            else:
                em.w(f"      S = fma(S, ir[j], t * q{qmins[n]});")
        em.w("      s[o * NPT + j] = S;")
        em.w("    }")

    emit_blocked_sum(em, "q_blocked", monos, groups, combine)


def emit_r_kappa(em: Emitter, records, ks_list, n_ix):
    """Contracted R-family, far field: the same factoring as the Q-family with
    depth-dependent record coefficients C_rec(z) = sum_pz C[rec, pz] z^pz (z
    the receiver's x3, fixed per block).  Per quadrature point on the mirrored
    triangle sum_k kappa_k d1^a d2^b d3^c R^-n, n odd, combined by Horner in
    1/R^2."""
    by_mono = defaultdict(list)
    for rec, ks in zip(records, ks_list):
        by_mono[tuple(rec[n_ix:n_ix + 3]) + (rec[n_ix + 3],)].append((slot_of(rec[:n_ix]), ks))
    monos = _walk_order(list(by_mono))
    groups = sorted({m[3] for m in monos})
    assert groups == list(range(groups[0], groups[-1] + 1, 2))
    kidx = {m: i for i, m in enumerate(monos)}
    keyed = []
    for m in monos:
        for slot, ks in by_mono[m]:
            i, j, p, mm = slot // 27, (slot // 9) % 3, (slot // 3) % 3, slot % 3
            for pz, ci in ks.items():
                keyed.append((kidx[m] * 6 + SYM[(j, mm)], ci, pz, i * 3 + p))
    em.w(f"#define MHS_RM {len(monos)}")
    em.w(f"#define MHS_RREC {len(keyed)}")
    em.w("// R B-table CSR over (monomial k, symmetric source index v): coefficient")
    em.w("// index, power of the receiver depth, receiver index i * 3 + p.")
    _csr(em, "mhs_rb", keyed, len(monos) * 6, [("coef", "short"), ("pz", "unsigned char"),
                                                 ("ip", "unsigned char")])

    def combine(em, gidx):
        em.w("    const T i2 = ir[j] * ir[j];")
        em.w(f"    const T i{groups[0]} = {' * '.join(['ir[j]'] * groups[0])};")
        em.w("#pragma unroll")
        em.w("    for (int o = 0; o < NOUT; ++o) {")
        em.w(f"      T S = acc[o][j][{gidx[groups[-1]]}];")
        for n in reversed(groups[:-1]):
            em.w(f"      S = fma(S, i2, acc[o][j][{gidx[n]}]);")
        em.w(f"      s[o * NPT + j] = S * i{groups[0]};")
        em.w("    }")

    emit_blocked_sum(em, "r_blocked", monos, groups, combine)


def emit_q_moments(em: Emitter, fam, records, n_ix, base, chunk):
    """Moments-first Q-family for the uncontracted forms: per quadrature point
    only the monomial moments M_k += w d1^a d2^b d3^c R^-n Q^-q of one chunk
    of monomials; per pair each chunk's records are then folded into the
    traction-operator tensor S (image.py:107-162): for a DDG record
    (i, j, p, m) of value v = C M_k, S[i][p][j] += mu n_m v,
    S[i][p][m] += mu n_j v and, if j == m, S[i][p][k] += lam n_k v; for a DG
    record (i, j, m) the same with the p index dropped."""
    by_mono = defaultdict(list)
    for r, rec in enumerate(records):
        by_mono[tuple(rec[n_ix:n_ix + 5])].append((rec[:n_ix], base + r))
    monos = sorted(by_mono, key=lambda m: (m[:3], m[3:]))
    chunks = [monos[i:i + chunk] for i in range(0, len(monos), chunk)]
    em.w(f"#define MHS_QMOM_{fam.upper()}_CHUNKS {len(chunks)}")
    em.w(f"#define MHS_QMOM_{fam.upper()}_CHUNK {chunk}")
    for c, ms in enumerate(chunks):
        # accumulate the chunk's moments
        em.w(f"__device__ __forceinline__ void qmom_{fam}_acc{c}(double x, double y, double z, double ir,")
        em.w("    double iq, double w, double* M)")
        em.w("{")
        amax = max(m[0] for m in ms); bmax = max(m[1] for m in ms); cmax = max(m[2] for m in ms)
        nmax = max(m[3] for m in ms); qmax = max(m[4] for m in ms)
        for nm, var, e in (("x", "x", amax), ("y", "y", bmax), ("z", "z", cmax), ("r", "ir", nmax),
                           ("q", "iq", qmax)):
            if e >= 1:
                em.w(f"  const double {nm}1 = {var};")
            for k in range(2, e + 1):
                em.w(f"  const double {nm}{k} = {nm}{k - 1} * {var};")
        seen_d, seen_r = set(), set()
        for j, (a, b, cz, n, q) in enumerate(ms):
            dn = f"D{a}{b}{cz}"
            if (a, b, cz) not in seen_d:
                f = [f"x{a}"] * (a > 0) + [f"y{b}"] * (b > 0) + [f"z{cz}"] * (cz > 0)
                em.w(f"  const double {dn} = {' * '.join(f) if f else '1.0'};")
                seen_d.add((a, b, cz))
            rn = f"R{n}_{q}"
            if (n, q) not in seen_r:
                f = [f"r{n}"] * (n > 0) + [f"q{q}"] * (q > 0)
                em.w(f"  const double {rn} = w * {' * '.join(f) if f else '1.0'};")
                seen_r.add((n, q))
            em.w(f"  M[{j}] = fma({dn}, {rn}, M[{j}]);")
        em.w("}")
        em.w()
        # fold the chunk's records into S
        em.w(f"__device__ __forceinline__ void qmom_{fam}_fold{c}(const double* M, const double* __restrict__ C,")
        em.w("    const double* n, double mu, double lam, double* S)")
        em.w("{")
        em.w("  const double mn0 = mu * n[0], mn1 = mu * n[1], mn2 = mu * n[2];")
        em.w("  const double ln0 = lam * n[0], ln1 = lam * n[1], ln2 = lam * n[2];")
        for j, m in enumerate(ms):
            for idx, ci in by_mono[m]:
                if n_ix == 4:
                    i, jj, p, mm = idx
                    base_ix = (i * 3 + p) * 3
                else:
                    i, jj, mm = idx
                    base_ix = i * 3
                em.w(f"  {{ const double v = C[{ci}] * M[{j}];")
                em.w(f"    S[{base_ix + jj}] = fma(mn{mm}, v, S[{base_ix + jj}]);")
                em.w(f"    S[{base_ix + mm}] = fma(mn{jj}, v, S[{base_ix + mm}]);")
                if jj == mm:
                    for k in range(3):
                        em.w(f"    S[{base_ix + k}] = fma(ln{k}, v, S[{base_ix + k}]);")
                em.w("  }")
        em.w("}")
        em.w()
    # dispatchers
    for kind in ("acc", "fold"):
        if kind == "acc":
            em.w(f"__device__ __forceinline__ void qmom_{fam}_acc(int c, double x, double y, double z, double ir,")
            em.w("    double iq, double w, double* M)")
        else:
            em.w(f"__device__ __forceinline__ void qmom_{fam}_fold(int c, const double* M, const double* __restrict__ C,")
            em.w("    const double* n, double mu, double lam, double* S)")
        em.w("{")
        em.w("  switch (c) {")
        for c in range(len(chunks)):
            args = "x, y, z, ir, iq, w, M" if kind == "acc" else "M, C, n, mu, lam, S"
            em.w(f"    case {c}: qmom_{fam}_{kind}{c}({args}); break;")
        em.w("    default: break;")
        em.w("  }")
        em.w("}")
        em.w()


def main():
    t = json.loads(TABLE.read_text())
    em = Emitter()
    em.w("// GENERATED by mhs_amd/gen/emit.py from mhs_amd/data/mindlin_table.json -- do not edit.")
    em.w("#pragma once")
    em.w()
    body = Emitter()
    body.coef_src = em.coef_src
    offs = {}
    offs["r_ddg"] = body.n_coef
    emit_r(body, "r_ddg", t["r_ddg"], 4, 81)
    offs["r_dg"] = body.n_coef = len(body.coef_src)
    emit_r(body, "r_dg", t["r_dg"], 3, 27)
    offs["q_ddg"] = body.n_coef = len(body.coef_src)
    body.q_ddg_base = offs["q_ddg"]
    emit_q(body, "q_ddg_point", t["q_ddg"], 4)
    offs["q_dg"] = body.n_coef = len(body.coef_src)
    emit_q(body, "q_dg_point", t["q_dg"], 3)
    n_coef = len(body.coef_src)
    em.w(f"#define MHS_NCOEF {n_coef}")
    for k, v in offs.items():
        em.w(f"#define MHS_COEF_{k.upper()} {v}")
    em.w()
    # coefficient data: for each coefficient a list of (pnu, num, den)
    flat = []
    starts = [0]
    for terms in body.coef_src:
        for pnu, fr in terms:
            flat.append((pnu, fr.numerator, fr.denominator))
        starts.append(len(flat))
    em.w(f"static const int mhs_coef_start[{len(starts)}] = {{{', '.join(map(str, starts))}}};")
    em.w(f"static const double mhs_coef_terms[{len(flat)}][3] = {{")
    for pnu, num, den in flat:
        em.w(f"  {{{pnu}, {num}.0, {den}.0}},")
    em.w("};")
    em.w()
    em.w("// Per-material coefficients: each record's sum num/den nu^pnu times the")
    em.w("// common scale 2 (lam + mu) / (pi mu (lam + 2 mu)).  The R offsets index")
    em.w("// the DDG block first, then DG; the Q offsets likewise.")
    em.w("static inline void mhs_fill_coeffs(double mu, double lam, double* c)")
    em.w("{")
    em.w("  const double nu = lam / (2.0 * (lam + mu));")
    em.w("  const double scale = 2.0 * (lam + mu) / (3.14159265358979323846 * mu * (lam + 2.0 * mu));")
    em.w(f"  for (int k = 0; k < {n_coef}; ++k) {{")
    em.w("    double s = 0.0;")
    em.w("    for (int t = mhs_coef_start[k]; t < mhs_coef_start[k + 1]; ++t) {")
    em.w("      double pw = 1.0;")
    em.w("      for (int e = 0; e < (int)mhs_coef_terms[t][0]; ++e) pw *= nu;")
    em.w("      s += (mhs_coef_terms[t][1] / mhs_coef_terms[t][2]) * pw;")
    em.w("    }")
    em.w("    c[k] = s * scale;")
    em.w("  }")
    em.w("}")
    em.w()
    em.lines += body.lines
    kap = Emitter()
    kap.q_ddg_base = offs["q_ddg"]
    emit_q_kappa(kap, t["q_ddg"], 4)
    emit_r_kappa(kap, t["r_ddg"], body.rec_ks["r_ddg"], 4)
    emit_q_moments(kap, "ddg", t["q_ddg"], 4, offs["q_ddg"], 42)
    emit_q_moments(kap, "dg", t["q_dg"], 3, offs["q_dg"], 35)
    em.lines += kap.lines
    OUT.write_text("\n".join(em.lines) + "\n")
    print(f"wrote {OUT} ({len(em.lines)} lines, {n_coef} coefficients)")


if __name__ == "__main__":
    main()
