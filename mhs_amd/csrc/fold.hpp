// The Boussinesq correction of a folded (surface-reaching) source: the
// receiver-centred polar surface grid and the point-load kernels.
//
// Reference: src/mhs/kernels/fold.py:95-367, constants defaults.py:108-148.
#pragma once
#include "pair.hpp"

#define MHS_PI 3.14159265358979323846
#define FOLD_MAXB 192
#define FOLD_MAXP (FOLD_MAXB - 1)

// Boussinesq unit compressive point load: stress components
// (xx, yy, xy, xz, yz, zz) at horizontal offset (dx, dy) = receiver - load and
// depth zeta (fold.py:132-166).
MHS_INLINE void point_stress(double dx, double dy, double zeta, double nu, double S[6]) {
  const double r2 = dx * dx + dy * dy;
  const double rho = sqrt(r2 + zeta * zeta);
  const double rho3 = rho * rho * rho;
  const double rho5 = rho3 * rho * rho;
  const double a = 1.0 / (rho * (rho + zeta));
  const double c = 1.0 - 2.0 * nu;
  const double s_th = c * (-a + zeta / rho3);
  const double s_rth = 2.0 * c * a - 3.0 * zeta * r2 / rho5 - c * zeta / rho3;
  const double r = sqrt(r2);
  const double ex = r > 0.0 ? dx / r : 0.0;
  const double ey = r > 0.0 ? dy / r : 0.0;
  const double pref = 1.0 / (2.0 * MHS_PI);
  S[0] = pref * (s_th + s_rth * ex * ex);
  S[1] = pref * (s_th + s_rth * ey * ey);
  S[2] = pref * s_rth * ex * ey;
  const double t_rz = pref * 3.0 * r * zeta * zeta / rho5;
  S[3] = t_rz * ex;
  S[4] = t_rz * ey;
  S[5] = -pref * 3.0 * zeta * zeta * zeta / rho5;
}

MHS_INLINE void point_disp(double dx, double dy, double zeta, double mu, double nu, double U[3]) {
  const double rho = sqrt(dx * dx + dy * dy + zeta * zeta);
  const double rho3 = rho * rho * rho;
  const double pref = 1.0 / (4.0 * MHS_PI * mu);
  const double ur = pref * (zeta / rho3 - (1.0 - 2.0 * nu) / (rho * (rho + zeta)));
  U[0] = ur * dx;
  U[1] = ur * dy;
  U[2] = -pref * (2.0 * (1.0 - nu) / rho + zeta * zeta / rho3);
}

// Love's on-axis uniform disk (fold.py:185-210): (srr, szz) and u_z.
MHS_INLINE void disk_stress(double a, double zeta, double nu, double& srr, double& szz) {
  const double rho = sqrt(a * a + zeta * zeta);
  const double q = zeta * zeta * zeta / (rho * rho * rho);
  szz = -(1.0 - q);
  srr = -0.5 * ((1.0 + 2.0 * nu) - 2.0 * (1.0 + nu) * zeta / rho + q);
}

MHS_INLINE double disk_uz(double a, double zeta, double mu, double nu) {
  const double rho = sqrt(a * a + zeta * zeta);
  return -(0.5 / mu) * (2.0 * (1.0 - nu) * (rho - zeta) + zeta - zeta * zeta / rho);
}

MHS_INLINE double halfwidth(double ra, double rb, double rho_s, double radius) {
  if (rb < rho_s - radius || ra > rho_s + radius) return 0.0;
  if (rho_s <= radius) return MHS_PI;
  return asin(fmin(1.0, radius / rho_s));
}

// Polar-grid layout of one fold pair, built by one thread (fold.py:290-330).
struct FoldGrid {
  int np;                        // panels
  int total;                     // points
  double R1, rho_s, theta_c;
  double ra[FOLD_MAXP], rb[FOLD_MAXP];
  double t0[FOLD_MAXP][4], t1[FOLD_MAXP][4];
  int m[FOLD_MAXP][4];
  int npc[FOLD_MAXP];            // pieces
  int cum[FOLD_MAXP + 1];        // cumulative point counts
  int overflow;
};

// One folded pair's slot of the fold work buffer: its polar grid and the
// surface load at the receiver's foot for every node, prepared in parallel
// (a thread per pair) before the point sums.
struct FoldItem {
  FoldGrid G;
  double s0[6][3];
};

// numpy linspace node i of (start, stop, num): i * step + start, stop exact.
MHS_INLINE double lin_node(double start, double stop, int num, int i) {
  if (i == num - 1) return stop;
  const double step = (stop - start) / (num - 1);
  return i * step + start;
}

MHS_INLINE void fold_grid_build(FoldGrid& G, double xh0, double xh1, double zeta, double cx, double cy,
                                double r_f, double eps, double diam) {
  G.overflow = 0;
  const double dxc = cx - xh0, dyc = cy - xh1;
  const double rho_s = hypot(dxc, dyc);
  const double theta_c = atan2(dyc, dxc);
  G.rho_s = rho_s;
  G.theta_c = theta_c;
  const double r1_target = rho_s + 4 * r_f + 8.0 * diam;
  double geo[64];
  int ng = 0;
  geo[ng++] = 0.0;
  if (0.0 < zeta && zeta < 0.5 * eps) {
    const double zz[4] = {0.5 * zeta, zeta, 2 * zeta, 4 * zeta};
    for (int i = 0; i < 4; ++i)
      if (zz[i] < 0.5 * eps) geo[ng++] = zz[i];
  }
  double r = 0.5 * eps;
  while (true) {
    geo[ng++] = r;
    if (r >= r1_target || ng >= 63) break;
    r *= 2.0;
  }
  const double R1 = geo[ng - 1];
  G.R1 = R1;
  const double lo_m = rho_s - 4 * r_f, hi_m = rho_s + 4 * r_f;
  const double lo_f = rho_s - r_f, hi_f = rho_s + r_f;
  double brk[FOLD_MAXB + 8];
  int nb = 0;
  for (int i = 0; i < ng; ++i)
    if (geo[i] < lo_m || geo[i] > hi_m) brk[nb++] = geo[i];
  const double A0[3] = {fmax(0.0, lo_m), fmax(0.0, lo_f), fmin(R1, hi_f)};
  const double B0[3] = {fmax(0.0, lo_f), fmin(R1, hi_f), fmin(R1, hi_m)};
  const double WD[3] = {16.0, 6.0, 16.0};
  for (int s = 0; s < 3; ++s) {
    if (!(B0[s] > A0[s])) continue;
    const int num = (int)fmax(1.0, ceil((B0[s] - A0[s]) / (WD[s] * eps))) + 1;
    for (int i = 0; i < num; ++i) {
      if (nb >= FOLD_MAXB) { G.overflow = 1; break; }
      brk[nb++] = lin_node(A0[s], B0[s], num, i);
    }
  }
  // sort + unique + clip at R1 (np.unique, fold.py:317-318)
  for (int i = 1; i < nb; ++i) {
    const double v = brk[i];
    int j = i - 1;
    while (j >= 0 && brk[j] > v) { brk[j + 1] = brk[j]; --j; }
    brk[j + 1] = v;
  }
  int nu_ = 0;
  for (int i = 0; i < nb; ++i) {
    if (nu_ > 0 && brk[i] == brk[nu_ - 1]) continue;
    if (!(brk[i] <= R1 + 1e-12 * R1)) continue;
    brk[nu_++] = brk[i];
  }
  int np = 0;
  G.cum[0] = 0;
  const double n = 4.0;
  for (int i = 0; i + 1 < nu_; ++i) {
    const double ra = brk[i], rb = brk[i + 1];
    if (rb - ra < 1e-9 * eps) continue;
    if (np >= FOLD_MAXP) { G.overflow = 1; break; }
    const double wf = halfwidth(ra, rb, rho_s, r_f);
    const double wm = halfwidth(ra, rb, rho_s, 4 * r_f);
    int npc = 0;
    double ts[4], te[4];
    int mm[4];
    if (wf >= MHS_PI) {
      ts[0] = 0.0; te[0] = 2 * MHS_PI;
      mm[0] = (int)fmax(8.0, ceil(2 * MHS_PI * rb / (n * 0.25 * eps)));
      npc = 1;
    } else if (wm >= MHS_PI) {
      ts[0] = 0.0; te[0] = 2 * MHS_PI;
      mm[0] = (int)fmax(8.0, ceil(2 * MHS_PI * rb / (n * 2.0 * eps)));
      npc = 1;
    } else {
      // order: medium-left, fine, medium-right, coarse (fold.py:270-287)
      const bool has_f = wf > 0.0, has_m = wm > wf;
      if (has_m) {
        ts[npc] = theta_c - wm; te[npc] = theta_c - wf;
        mm[npc] = (int)fmax(1.0, ceil((wm - wf) * rb / (n * 2.0 * eps)));
        ++npc;
      }
      if (has_f) {
        ts[npc] = theta_c - wf; te[npc] = theta_c + wf;
        mm[npc] = (int)fmax(1.0, ceil(2 * wf * rb / (n * 0.25 * eps)));
        ++npc;
      }
      if (has_m) {
        ts[npc] = theta_c + wf; te[npc] = theta_c + wm;
        mm[npc] = (int)fmax(1.0, ceil((wm - wf) * rb / (n * 2.0 * eps)));
        ++npc;
      }
      const double start = wm > 0.0 ? theta_c + wm : 0.0;
      ts[npc] = start; te[npc] = start + 2 * MHS_PI - 2 * wm;
      mm[npc] = 6;                   // max(1, FOLD_N_COARSE // 4)
      ++npc;
    }
    int cnt = 0, kept = 0;
    for (int q = 0; q < npc; ++q) {
      if (!(te[q] > ts[q])) continue;
      G.t0[np][kept] = ts[q]; G.t1[np][kept] = te[q]; G.m[np][kept] = mm[q];
      cnt += 4 * mm[q];
      ++kept;
    }
    G.npc[np] = kept;
    G.ra[np] = ra; G.rb[np] = rb;
    G.cum[np + 1] = G.cum[np] + 6 * cnt;
    ++np;
  }
  G.np = np;
  G.total = G.cum[np];
}

// Node and weight of flat point index g of the grid: returns (x, y, w).
MHS_INLINE void fold_grid_point(const FoldGrid& G, double xh0, double xh1, const double* u6, const double* w6,
                                const double* u4, const double* w4, int g, double& x, double& y, double& w) {
  int lo = 0, hi = G.np;                         // cum[lo] <= g < cum[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (G.cum[mid] <= g) lo = mid; else hi = mid;
  }
  const int p = lo;
  int li = g - G.cum[p];
  const int nang = (G.cum[p + 1] - G.cum[p]) / 6;
  const int ir = li / nang;
  int a = li - ir * nang;
  const double ra = G.ra[p], rb = G.rb[p];
  const double r = ra + (rb - ra) * u6[ir];
  const double wr = (rb - ra) * w6[ir];
  int q = 0;
  while (a >= 4 * G.m[p][q]) { a -= 4 * G.m[p][q]; ++q; }
  const int sg = a >> 2, j = a & 3;
  const int num = G.m[p][q] + 1;
  const double tlo = lin_node(G.t0[p][q], G.t1[p][q], num, sg);
  const double thi = lin_node(G.t0[p][q], G.t1[p][q], num, sg + 1);
  const double th = tlo + (thi - tlo) * u4[j];
  const double wt = (thi - tlo) * w4[j];
  x = xh0 + r * cos(th);
  y = xh1 + r * sin(th);
  w = wr * wt * r;
}

// The surface load's contraction, split into what a fold item fixes and
// what a surface point adds.  direct_contract with the receiver half
// E = half(e_z, e_z) and the source half F_m = half(e_m, n) (slip along axis
// m, n the source normal) forms the quadratic forms d.A.d of E, F_m and
// sym(E F_m) in the source frame and its scalar coefficients from those
// halves alone; only the forms' z terms and the tables W depend on the
// point.  FoldContract holds the point-independent part for the three slip
// directions, formed once per item; fold_contract then costs a point about
// a third of direct_contract, with the same expressions in the same order.
struct FoldContract {
  double L[7][6];     // E, F_0..F_2, sym(E F_0)..sym(E F_2): L00, L01, L11, L02, L12, L22 in the frame
  double A3[3], A5a[3], A5b[3], A5e[3], rFn[3];
  double A5c, A7c, C1, eigK;
};

MHS_INLINE void frame_form(const Frame& f, const double A[6], double* L) {
  const double M[3][3] = {{A[0], A[3], A[4]}, {A[3], A[1], A[5]}, {A[4], A[5], A[2]}};
  const double* Q[3] = {f.e1, f.e2, f.n};
  double R[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a)
        v += Q[i][a] * (M[a][0] * Q[j][0] + M[a][1] * Q[j][1] + M[a][2] * Q[j][2]);
      R[i][j] = v;
    }
  L[0] = R[0][0]; L[1] = R[0][1]; L[2] = R[1][1]; L[3] = R[0][2]; L[4] = R[1][2]; L[5] = R[2][2];
}

// local_quadratic's table from a form in the frame, at plane height z.
MHS_INLINE void form_at(const double* L, double z, double p[6]) {
  p[tix(0, 0)] = z * z * L[5];
  p[tix(1, 0)] = -2.0 * z * L[3];
  p[tix(0, 1)] = -2.0 * z * L[4];
  p[tix(2, 0)] = L[0];
  p[tix(1, 1)] = 2.0 * L[1];
  p[tix(0, 2)] = L[2];
}

MHS_INLINE void fold_contract_prep(const Params& pr, const Frame& f, const double* __restrict__ src,
                                   FoldContract& c) {
  const double nu = pr.nu;
  const double c34 = 3.0 - 4.0 * nu;
  const double eps = src[S_EPS];
  const double e2 = eps * eps;
  const double ez[3] = {0.0, 0.0, 1.0};
  double E[6];
  half_tensor(ez, ez, pr.mu, pr.lam, E);
  frame_form(f, E, c.L[0]);
  const double trE = E[0] + E[1] + E[2];
  const double Em[3][3] = {{E[0], E[3], E[4]}, {E[3], E[1], E[5]}, {E[4], E[5], E[2]}};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double em[3] = {m == 0 ? 1.0 : 0.0, m == 1 ? 1.0 : 0.0, m == 2 ? 1.0 : 0.0};
    double F[6];
    half_tensor(em, src + S_N, pr.mu, pr.lam, F);
    const double Fm[3][3] = {{F[0], F[3], F[4]}, {F[3], F[1], F[5]}, {F[4], F[5], F[2]}};
    double EF = 0.0, Ms[6];
    double MM[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        MM[i][j] = Em[i][0] * Fm[0][j] + Em[i][1] * Fm[1][j] + Em[i][2] * Fm[2][j];
        EF += Em[i][j] * Fm[i][j];
      }
    Ms[0] = MM[0][0]; Ms[1] = MM[1][1]; Ms[2] = MM[2][2];
    Ms[3] = 0.5 * (MM[0][1] + MM[1][0]); Ms[4] = 0.5 * (MM[0][2] + MM[2][0]); Ms[5] = 0.5 * (MM[1][2] + MM[2][1]);
    const double trF = F[0] + F[1] + F[2];
    frame_form(f, F, c.L[1 + m]);
    frame_form(f, Ms, c.L[4 + m]);
    c.A3[m] = (1.0 - c34) * EF + trE * trF;
    c.A5a[m] = -3.0 * trF;
    c.A5b[m] = -3.0 * trE;
    c.A5e[m] = -6.0 * (1.0 - nu) * e2 * EF;
    c.rFn[m] = F[2];
  }
  c.eigK = 15.0 * e2 * e2 / (8.0 * 3.14159265358979323846);
  c.A5c = 3.0 * c34 - 9.0;
  c.A7c = 30.0 * (1.0 - nu) * e2;
  c.C1 = 1.0 / (16.0 * 3.14159265358979323846 * pr.mu * (1.0 - nu));
}

// s[m] = 2 direct_contract(f, W, z, eps, mt, E, F_m, true, F_m[2]) for the
// three slip directions, from the item's FoldContract.
MHS_INLINE void fold_contract(const FoldContract& c, const double W[6][21], double z, double s[3]) {
  const double* W3 = W[2];
  const double* W5 = W[3];
  const double* W7 = W[4];
  constexpr int U[6] = {0, 1, 0, 2, 1, 0}, V[6] = {0, 0, 1, 0, 1, 2};   // exponents of tix 0..5
  double pa[6];
  form_at(c.L[0], z, pa);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double pb[6], pc[6];
    form_at(c.L[1 + m], z, pb);
    form_at(c.L[4 + m], z, pc);
    double S = c.A3[m] * W3[0] + c.A5e[m] * W5[0];
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      S = fma(c.A5c * pc[t] + c.A5a[m] * pa[t] + c.A5b[m] * pb[t], W5[t], S);
      S = fma(c.A7c * pc[t], W7[t], S);
    }
#pragma unroll
    for (int t1 = 0; t1 < 6; ++t1)
#pragma unroll
      for (int t2 = 0; t2 < 6; ++t2)
        S = fma(15.0 * pa[t1] * pb[t2], W7[tix(U[t1] + U[t2], V[t1] + V[t2])], S);
    double out = -c.C1 * S;
    out -= c.eigK * W7[0] * c.rFn[m];        // the eigenstress term
    s[m] = 2.0 * out;
  }
}

// Surface load s[m] = 2 (H - E C)_zz of node k at (x, y, 0), per slip m
// (fold.py:214-228).
template <int P>
MHS_INLINE void surface_load(const Params& pr, const double* __restrict__ src, const FoldContract& fc, int k,
                             double x, double y, double s[3]) {
  const double o[3] = {x, y, 0.0};
  const double eps = src[S_EPS];
  Frame f;
  load_frame(src, false, f);
  double W[6][21];
  double X1, X2;
  const double z = weighted_tables<SpecH<P>, P>(f, src + S_V, k, o, eps, pr.gl_x, pr.gl_w, W, X1, X2);
  // sigma_zz (receiver half of e_z on the plane z = 0) per unit slip e_m,
  // minus the eigenstress E C_zz, by the contracted lift
  fold_contract(fc, W, z, s);
}


// The same surface load at a point known to be far from the triangle
// (sqrt(|x - c|^2 + eps^2) > 10 L): the collapsed Gauss tables only, so a
// kernel evaluating far points alone carries none of the closed form's
// registers.
template <int P>
MHS_INLINE void surface_load_far(const Params& pr, const double* __restrict__ src, const FoldContract& fc, int k,
                                 double x, double y, double s[3]) {
  const double o[3] = {x, y, 0.0};
  const double eps = src[S_EPS];
  Frame f;
  load_frame(src, false, f);
  const double d0 = o[0] - f.c[0], d1 = o[1] - f.c[1], d2 = o[2] - f.c[2];
  const double z = d0 * f.n[0] + d1 * f.n[1] + d2 * f.n[2];
  const double X1 = d0 * f.e1[0] + d1 * f.e1[1] + d2 * f.e1[2];
  const double X2 = d0 * f.e2[0] + d1 * f.e2[1] + d2 * f.e2[2];
  const double h2 = z * z + eps * eps;
  const int ng = centroid_distance(f.c, o, eps) > 40.0 * f.L ? 8 : 12;
  const int off = ng * (ng - 1) / 2;
  double W[6][21];
  weighted_far<SpecH<P>, P>(f, k, X1, X2, h2, ng, pr.gl_x + off, pr.gl_w + off, W);
  fold_contract(fc, W, z, s);
}
