// Per-pair evaluation: direct + image (or folded mirror) + eigenstress terms,
// the Q-family point sum, and the output epilogues.
//
// Reference: src/mhs/kernels/assemble.py:96-331, image.py:170-254, fold.py:84-118.
#pragma once
#include "core.hpp"
#include "gen_image.inc"

enum Kind { K_DISP = 0, K_STRESS = 1, K_EIGEN = 2, K_TRACTION = 3, K_INTERACTION = 4 };
enum Flag { F_FOLD = 1, F_SUBTRACT = 2, F_PARTS = 4, F_MIXED = 8 };

struct Params {
  const double* obs;     // n_obs x 3
  const double* nrm;     // n_obs x 3 (traction, interaction)
  const double* rb;      // n_obs x n_a x 3 (interaction)
  const double* sb;      // n_src x n_b x 3 (interaction) or null = identity
  const double* src;     // n_src x MHS_SRC_STRIDE
  const double* coef;    // MHS_NCOEF image coefficients for (mu, lam)
  const double* gl_x;    // Gauss-Legendre nodes on [0, 1], orders 1..192
  const double* gl_w;
  const double* gq;      // flattened collapsed Gauss rules: x1 | x2 | w, MHS_GQ_TOTAL each
  double* out;
  double* out2;
  int64_t n_obs, n_src, n_cols;
  int64_t obs_lo;        // first observer row of this launch
  int64_t out_lo;        // observer row held by the first row of out / out2
  int kind, flags, n_a, n_b;
  // the rows of rb / sb and the entries of out per (row, column): n_a, n_b
  // and n_a n_b, except in an output group of the contracted interaction
  // (n_a = 1, n_b <= 2 of the whole launch's na_tot x nb_tot)
  int na_tot, nb_tot;
  int far_tiles;         // column tiles per block of the contracted far pass
  double mu, lam, nu;
  // constants of the contracted direct series, computed once on the host so
  // that the far pass holds them in scalar registers: C1 = 1/(16 pi mu (1-nu)),
  // A5 coefficient -6(1-nu), A7 coefficient 30(1-nu), eigenstress scale
  // 15/(8 pi C1), and 1 - c34, 3 c34 - 9 with c34 = 3 - 4 nu
  double C1, k5e, k7c, k7e, k3, k5c;
  // near/far switch of the contracted interaction's passes, in units of L:
  // 10 (the reference's D* = 10 L), or the mixed rule's near ratio
  double near_r;
  double fold_near_r;    // the same switch for the fold's surface points
  Rule rule;             // the mixed mode's far-pass rule
  int mixed_rule;        // 1: the launch runs the mixed rule (mixed mode, contracted interaction)
  // per 256 element columns: bounding box of the sources' centroids and
  // mirrored centroids (lo xyz, hi xyz), their largest L, any straddling
  // source (1/0); or null
  const double* near_tiles;
};

// The launch's mixed rule, or null for the reference's rule.
MHS_INLINE const Rule* rule_of(const Params& pr) { return pr.mixed_rule ? &pr.rule : nullptr; }

// Fold classification of one (receiver, source) pair (fold.py:84-118): the
// source straddles and the receiver is shallower than 10 eps and within
// 60 eps of the crease centre.
MHS_INLINE bool in_fold_band(const double* __restrict__ src, const double o[3], int flags) {
  if (!(flags & F_FOLD) || src[S_STRADDLE] == 0.0) return false;
  const double eps = src[S_EPS];
  const double dx = o[0] - src[S_CREASE], dy = o[1] - src[S_CREASE + 1];
  return (-o[2] < 10.0 * eps) && (hypot(dx, dy) < 60.0 * eps);
}

MHS_INLINE void reflected_vertices(const double* __restrict__ src, double v[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v[3 * i] = src[S_V + 3 * i];
    v[3 * i + 1] = src[S_V + 3 * i + 1];
    v[3 * i + 2] = -src[S_V + 3 * i + 2];
  }
}

// Hooke tensor C:sym(slip (x) n) per unit slip j (assemble.py:123-128).
MHS_INLINE double hooke(const double* n, int a, int b, int j, double mu, double lam) {
  return (a == b ? lam * n[j] : 0.0) + mu * ((a == j ? n[b] : 0.0) + (b == j ? n[a] : 0.0));
}

// Total stress of node k at o (direct + image R-family, or the folded mirror
// in the band) and the blob weights of the eigenstress terms.
template <int P>
MHS_INLINE void pair_stress(const Params& pr, const double* __restrict__ src, int k, const double o[3],
                            bool band, bool want_H, double H[27], double& phi_d, double& phi_m) {
  const double eps = src[S_EPS];
  const Mat mt{pr.mu, pr.lam, pr.nu};
  double W[6][21];
  double X1, X2;
  {
    Frame f;
    load_frame(src, false, f);
    const double z = weighted_tables<SpecH<P>, P>(f, src + S_V, k, o, eps, pr.gl_x, pr.gl_w, W, X1, X2, pr.near_r, rule_of(pr));
    phi_d = direct_E(W, eps);
    if (want_H) direct_H(f, W, z, eps, mt, H);
  }
  phi_m = 0.0;
  Frame g;
  load_frame(src, true, g);
  double vi[9];
  reflected_vertices(src, vi);
  if (band) {
    // mirror element, full space, slip columns signed (-1, -1, +1) (fold.py:370-380)
    const double z = weighted_tables<SpecH<P>, P>(g, vi, k, o, eps, pr.gl_x, pr.gl_w, W, X1, X2, pr.near_r, rule_of(pr));
    phi_m = direct_E(W, eps);
    if (want_H) {
      double Hm[27];
      direct_H(g, W, z, eps, mt, Hm);
#pragma unroll
      for (int t = 0; t < 27; ++t) H[t] += ((t % 3) == 2 ? Hm[t] : -Hm[t]);
    }
  } else if (want_H) {
    // image R-family in closed form on the reflected triangle (image.py:435-497)
    const double zp = weighted_tables<SpecI<P>, P>(g, vi, k, o, eps, pr.gl_x, pr.gl_w, W, X1, X2, pr.near_r, rule_of(pr));
    const double al[3] = {-g.e1[0], -g.e1[1], -g.e1[2]};
    const double be[3] = {-g.e2[0], -g.e2[1], -g.e2[2]};
    const double ga[3] = {zp * g.n[0], zp * g.n[1], zp * g.n[2]};
    double DDG[81];
#pragma unroll
    for (int t = 0; t < 81; ++t) DDG[t] = 0.0;
    r_ddg(W[2], W[3], W[4], W[5], al, be, ga, o[2], pr.coef, DDG);
    const double nrm[3] = {src[S_N], src[S_N + 1], src[S_N + 2]};
    double Hi[27];
    contract_H(DDG, nrm, pr.mu, pr.lam, Hi);
#pragma unroll
    for (int t = 0; t < 27; ++t) H[t] += Hi[t];
  }
}

template <int P>
MHS_INLINE void pair_disp(const Params& pr, const double* __restrict__ src, int k, const double o[3],
                          bool band, double U[9]) {
  const double eps = src[S_EPS];
  const Mat mt{pr.mu, pr.lam, pr.nu};
  double W[6][21];
  double X1, X2;
  {
    Frame f;
    load_frame(src, false, f);
    const double z = weighted_tables<SpecU<P>, P>(f, src + S_V, k, o, eps, pr.gl_x, pr.gl_w, W, X1, X2, pr.near_r, rule_of(pr));
    direct_U(f, W, z, eps, mt, U);
  }
  Frame g;
  load_frame(src, true, g);
  double vi[9];
  reflected_vertices(src, vi);
  if (band) {
    const double z = weighted_tables<SpecU<P>, P>(g, vi, k, o, eps, pr.gl_x, pr.gl_w, W, X1, X2, pr.near_r, rule_of(pr));
    double Um[9];
    direct_U(g, W, z, eps, mt, Um);
#pragma unroll
    for (int t = 0; t < 9; ++t) U[t] += ((t % 3) == 2 ? Um[t] : -Um[t]);
  } else {
    const double zp = weighted_tables<SpecI<P>, P>(g, vi, k, o, eps, pr.gl_x, pr.gl_w, W, X1, X2, pr.near_r, rule_of(pr));
    const double al[3] = {-g.e1[0], -g.e1[1], -g.e1[2]};
    const double be[3] = {-g.e2[0], -g.e2[1], -g.e2[2]};
    const double ga[3] = {zp * g.n[0], zp * g.n[1], zp * g.n[2]};
    double DG[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) DG[t] = 0.0;
    r_dg(W[2], W[3], W[4], W[5], al, be, ga, o[2], pr.coef, DG);
    const double nrm[3] = {src[S_N], src[S_N + 1], src[S_N + 2]};
    double Ui[9];
    contract_U(DG, nrm, pr.mu, pr.lam, Ui);
#pragma unroll
    for (int t = 0; t < 9; ++t) U[t] += Ui[t];
  }
}

// Exact point-to-triangle distance (image.py:170-208).
MHS_INLINE double triangle_distance(const double o[3], const double* v) {
  const double a[3] = {v[0], v[1], v[2]};
  double ab[3], ac[3], ap[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) { ab[i] = v[3 + i] - a[i]; ac[i] = v[6 + i] - a[i]; ap[i] = o[i] - a[i]; }
  const double nx = ab[1] * ac[2] - ab[2] * ac[1];
  const double ny = ab[2] * ac[0] - ab[0] * ac[2];
  const double nz = ab[0] * ac[1] - ab[1] * ac[0];
  const double d20 = ap[0] * ab[0] + ap[1] * ab[1] + ap[2] * ab[2];
  const double d21 = ap[0] * ac[0] + ap[1] * ac[1] + ap[2] * ac[2];
  const double d00 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
  const double d01 = ab[0] * ac[0] + ab[1] * ac[1] + ab[2] * ac[2];
  const double d11 = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
  const double den = d00 * d11 - d01 * d01;
  const double vv = (d11 * d20 - d01 * d21) / den;
  const double ww = (d00 * d21 - d01 * d20) / den;
  const double uu = 1.0 - vv - ww;
  if (uu >= 0.0 && vv >= 0.0 && ww >= 0.0)
    return fabs(ap[0] * nx + ap[1] * ny + ap[2] * nz) / sqrt(nx * nx + ny * ny + nz * nz);
  double best = 1e300;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const double* p0 = v + 3 * e;
    const double* p1 = v + 3 * ((e + 1) % 3);
    double ev[3], d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) { ev[i] = p1[i] - p0[i]; d[i] = o[i] - p0[i]; }
    double t = (d[0] * ev[0] + d[1] * ev[1] + d[2] * ev[2]) / (ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2]);
    t = fmin(fmax(t, 0.0), 1.0);
    const double cx = o[0] - (p0[0] + t * ev[0]);
    const double cy = o[1] - (p0[1] + t * ev[1]);
    const double cz = o[2] - (p0[2] + t * ev[2]);
    best = fmin(best, sqrt(cx * cx + cy * cy + cz * cz));
  }
  return best;
}

// Q-family Gauss order of the pair (image.py:211-254).
template <int P>
MHS_INLINE int q_order(const double* __restrict__ src, const double o[3]) {
  double vi[9];
  reflected_vertices(src, vi);
  const double delta = triangle_distance(o, vi);
  const double eps = src[S_EPS];
  const double nq = ceil(10.0 * src[S_L] / sqrt(delta * delta + eps * eps)) + P;
  return (int)fmin(fmax(nq, 8.0), 192.0);
}

// The same order, settled without the exact distance when a bounding-sphere
// lower bound on it already puts 10 L / sqrt(delta^2 + eps^2) below the floor
// (with a relative margin far above rounding): every such pair gets 8.
template <int P>
MHS_INLINE int q_order_fast(const double* __restrict__ src, const double o[3]) {
  const double d0 = o[0] - src[S_IC], d1 = o[1] - src[S_IC + 1], d2 = o[2] - src[S_IC + 2];
  const double lb = fmax(sqrt(d0 * d0 + d1 * d1 + d2 * d2) - src[S_RC], 0.0);
  const double eps = src[S_EPS];
  const double x = 10.0 * src[S_L] / sqrt(lb * lb + eps * eps);
  if (x <= (8.0 - P) * (1.0 - 1e-9)) return 8;
  return q_order<P>(src, o);
}

// ------------------------------------------------------------- epilogues --
// Write (or add) the total stress H of one (observer row o, column col) into
// the output form of pr.kind.  eig terms: phi_d with the element normal nd,
// phi_m with the mirror normal nm and the mirror slip sign.
MHS_INLINE void emit_stress(const Params& pr, int64_t o, int64_t col, int64_t s, const double* H,
                            double phi_d, double phi_m, const double* nd, const double* nm,
                            bool add, bool add_eig = false) {
  double E[27];
  const bool eig = (!add || add_eig) && (pr.kind == K_EIGEN || (pr.flags & (F_SUBTRACT | F_PARTS)));
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double e = 0.0;
        if (eig) {
          e = phi_d * hooke(nd, a, b, j, pr.mu, pr.lam);
          if (phi_m != 0.0) e += phi_m * hooke(nm, a, b, j, pr.mu, pr.lam) * (j == 2 ? 1.0 : -1.0);
        }
        E[(a * 3 + b) * 3 + j] = e;
      }
  const int64_t oo = o - pr.out_lo;
  if (pr.kind == K_STRESS || pr.kind == K_EIGEN) {
    // The stress form keeps the total stress in out and the eigenstress in
    // out2; the elastic stress is formed afterwards as out - out2, once every
    // pass has added its part (matrices.py:287-294 subtracts the finished
    // eigenstress matrix from the finished total).  The eigenstress form
    // runs this same kernel and stores only E, so the three entry points
    // agree bitwise.
    const bool split = (pr.flags & (F_SUBTRACT | F_PARTS)) != 0;
#pragma unroll
    for (int ab = 0; ab < 9; ++ab)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int64_t idx = ((oo * 9 + ab) * pr.n_cols + col) * 3 + j;
        const double val = (pr.kind == K_EIGEN) ? E[ab * 3 + j] : H[ab * 3 + j];
        if (add) pr.out[idx] += val; else pr.out[idx] = val;
        if (!add && split && pr.kind == K_STRESS) pr.out2[idx] = E[ab * 3 + j];
      }
    return;
  }
  // traction t[i][j] = sum_b (H - E)[i][b][j] n_b
  const double* n = pr.nrm + 3 * o;
  const bool sub = (pr.flags & F_SUBTRACT) != 0;
  double t[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int b = 0; b < 3; ++b) acc += (H[(i * 3 + b) * 3 + j] - (sub ? E[(i * 3 + b) * 3 + j] : 0.0)) * n[b];
      t[i * 3 + j] = acc;
    }
  if (pr.kind == K_TRACTION) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int64_t idx = ((oo * 3 + i) * pr.n_cols + col) * 3 + j;
        if (add) pr.out[idx] += t[i * 3 + j]; else pr.out[idx] = t[i * 3 + j];
      }
    return;
  }
  // interaction: out[o, col, a, b] = rb[o, a, :] . t . sb[s, b, :]
  const double* rb = pr.rb + (int64_t)3 * pr.n_a * o;
  for (int a = 0; a < pr.n_a; ++a) {
    double ta[3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
      ta[j] = rb[3 * a] * t[j] + rb[3 * a + 1] * t[3 + j] + rb[3 * a + 2] * t[6 + j];
    for (int b = 0; b < pr.n_b; ++b) {
      double v;
      if (pr.sb) {
        const double* sb = pr.sb + ((int64_t)s * pr.n_b + b) * 3;
        v = ta[0] * sb[0] + ta[1] * sb[1] + ta[2] * sb[2];
      } else {
        v = ta[b];
      }
      const int64_t idx = ((oo * pr.n_cols + col) * pr.n_a + a) * pr.n_b + b;
      if (add) pr.out[idx] += v; else pr.out[idx] = v;
    }
  }
}

MHS_INLINE void emit_disp(const Params& pr, int64_t o, int64_t col, const double* U, bool add) {
  const int64_t oo = o - pr.out_lo;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int64_t idx = ((oo * 3 + i) * pr.n_cols + col) * 3 + j;
      if (add) pr.out[idx] += U[i * 3 + j]; else pr.out[idx] = U[i * 3 + j];
    }
}

// ------------------------------------------------------------- work items --
// One (observer row o, element column col) of the pair pass.
template <int P, int MODE>   // MODE 0: stress family (eigenstress included), 1: displacement
MHS_INLINE void pair_item(const Params& pr, int64_t o, int64_t col) {
  constexpr int K = (P + 1) * (P + 2) / 2;
  const int64_t s = col / K;
  const int k = (int)(col - s * K);
  const double* src = pr.src + s * MHS_SRC_STRIDE;
  const double ob[3] = {pr.obs[3 * o], pr.obs[3 * o + 1], pr.obs[3 * o + 2]};
  const bool band = in_fold_band(src, ob, pr.flags);
  if (MODE == 1) {
    double U[9];
    pair_disp<P>(pr, src, k, ob, band, U);
    emit_disp(pr, o, col, U, false);
  } else {
    double H[27];
    double phi_d, phi_m;
    pair_stress<P>(pr, src, k, ob, band, true, H, phi_d, phi_m);
    emit_stress(pr, o, col, s, H, phi_d, phi_m, src + S_N, src + S_IN, false);
  }
}

// Which near parts a pair of the contracted interaction has, by the same
// tests as pair_near_item: bit 0 the closed-form direct term, bit 1 the
// folded mirror (band) or the closed-form image R-family.
MHS_INLINE int near_parts(const Params& pr, const double* __restrict__ src, const double ob[3]) {
  const double eps = src[S_EPS];
  const bool band = in_fold_band(src, ob, pr.flags);
  const bool dnear = centroid_distance(src + S_C, ob, eps) <= pr.near_r * src[S_L];
  const bool rnear = !band && centroid_distance(src + S_IC, ob, eps) <= pr.near_r * src[S_IL];
  return (dnear ? 1 : 0) | ((band || rnear) ? 2 : 0);
}

// Near part of one (observer row o, element column col) of a contracted
// interaction, added after the far pass: the closed-form direct term and its
// eigenstress when the pair is near (sqrt(|x - c|^2 + eps^2) <= 10 L), and on
// the image side either the folded mirror element (fold band, with its
// eigenstress) or the closed-form R-family when the mirrored triangle is
// near.  Everything far is the far pass's.
template <int P, int NM>      // NM: the largest receiver / source component count (2 or 3)
MHS_INLINE void pair_near_item(const Params& pr, int64_t o, int64_t col) {
  constexpr int K = (P + 1) * (P + 2) / 2;
  const int64_t s = col / K;
  const int k = (int)(col - s * K);
  const double* src = pr.src + s * MHS_SRC_STRIDE;
  const double ob[3] = {pr.obs[3 * o], pr.obs[3 * o + 1], pr.obs[3 * o + 2]};
  const double eps = src[S_EPS];
  const bool band = in_fold_band(src, ob, pr.flags);
  const bool dnear = centroid_distance(src + S_C, ob, eps) <= pr.near_r * src[S_L];
  const bool rnear = !band && centroid_distance(src + S_IC, ob, eps) <= pr.near_r * src[S_IL];
  if (!dnear && !band && !rnear) return;
  const Mat mt{pr.mu, pr.lam, pr.nu};
  const bool sub = (pr.flags & F_SUBTRACT) != 0;
  // every output of the launch (up to NM x NM): the weighted tables, the
  // costly part, are formed once and contracted with each (a, b)
  const int n_a = pr.n_a, n_b = pr.n_b;
  const double* nr = pr.nrm + 3 * o;
  // receiver halves E_a, source halves F_b (element) and G_b (folded mirror:
  // slip (-sx, -sy, sz) on the mirror normal, fold.py:370-380)
  double Ea[NM][6], Fb[NM][6], Gb[NM][6], rF[NM][NM], rG[NM][NM];
#pragma unroll
  for (int a = 0; a < NM; ++a)
    if (a < n_a) half_tensor(pr.rb + ((int64_t)o * pr.na_tot + a) * 3, nr, pr.mu, pr.lam, Ea[a]);
#pragma unroll
  for (int b = 0; b < NM; ++b) {
    if (b >= n_b) continue;
    const double* sb = pr.sb + (s * pr.nb_tot + b) * 3;
    half_tensor(sb, src + S_N, pr.mu, pr.lam, Fb[b]);
    const double sm[3] = {-sb[0], -sb[1], sb[2]};
    half_tensor(sm, src + S_IN, pr.mu, pr.lam, Gb[b]);
#pragma unroll
    for (int a = 0; a < NM; ++a) {
      if (a >= n_a) continue;
      const double* rb = pr.rb + ((int64_t)o * pr.na_tot + a) * 3;
      double f = 0.0, g = 0.0;
      const int ix[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          f += rb[i] * Fb[b][ix[i][j]] * nr[j];
          g += rb[i] * Gb[b][ix[i][j]] * nr[j];
        }
      rF[a][b] = f;
      rG[a][b] = g;
    }
  }
  double acc[NM * NM];
#pragma unroll
  for (int t = 0; t < NM * NM; ++t) acc[t] = 0.0;
  double W[6][21];
  double X1, X2;
  if (dnear) {
    Frame f;
    load_frame(src, false, f);
    const double z = weighted_tables<SpecH<P>, P>(f, src + S_V, k, ob, eps, pr.gl_x, pr.gl_w, W, X1, X2, pr.near_r, rule_of(pr));
#pragma unroll
    for (int a = 0; a < NM; ++a)
#pragma unroll
      for (int b = 0; b < NM; ++b)
        if (a < n_a && b < n_b) acc[a * NM + b] += direct_contract(f, W, z, eps, mt, Ea[a], Fb[b], sub, rF[a][b]);
  }
  if (band || rnear) {
    Frame g;
    load_frame(src, true, g);
    double vi[9];
    reflected_vertices(src, vi);
    if (band) {
      const double z = weighted_tables<SpecH<P>, P>(g, vi, k, ob, eps, pr.gl_x, pr.gl_w, W, X1, X2, pr.near_r, rule_of(pr));
#pragma unroll
      for (int a = 0; a < NM; ++a)
#pragma unroll
        for (int b = 0; b < NM; ++b)
          if (a < n_a && b < n_b) acc[a * NM + b] += direct_contract(g, W, z, eps, mt, Ea[a], Gb[b], sub, rG[a][b]);
    } else {
      const double zp = weighted_tables<SpecI<P>, P>(g, vi, k, ob, eps, pr.gl_x, pr.gl_w, W, X1, X2, pr.near_r, rule_of(pr));
      const double al[3] = {-g.e1[0], -g.e1[1], -g.e1[2]};
      const double be[3] = {-g.e2[0], -g.e2[1], -g.e2[2]};
      const double ga[3] = {zp * g.n[0], zp * g.n[1], zp * g.n[2]};
      double DDG[81];
#pragma unroll
      for (int t = 0; t < 81; ++t) DDG[t] = 0.0;
      r_ddg(W[2], W[3], W[4], W[5], al, be, ga, ob[2], pr.coef, DDG);
      // rb.H.nr.sb of the R-family = sum over slots E[i][p] F[j][m] DDG[i][j][p][m]
      const int ix[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
#pragma unroll
      for (int a = 0; a < NM; ++a)
#pragma unroll
        for (int b = 0; b < NM; ++b) {
          if (a >= n_a || b >= n_b) continue;
          double v = 0.0;
#pragma unroll
          for (int t = 0; t < 81; ++t) {
            const int i = t / 27, j = (t / 9) % 3, p = (t / 3) % 3, m = t % 3;
            v = fma(Ea[a][ix[i][p]] * Fb[b][ix[j][m]], DDG[t], v);
          }
          acc[a * NM + b] += v;
        }
    }
  }
  const int64_t oo = o - pr.out_lo;
  const int os = pr.na_tot * pr.nb_tot;     // entries per (row, column)
#pragma unroll
  for (int a = 0; a < NM; ++a)
#pragma unroll
    for (int b = 0; b < NM; ++b)
      if (a < n_a && b < n_b) pr.out[(oo * pr.n_cols + col) * os + a * pr.nb_tot + b] += acc[a * NM + b];
}

