// Device math of the mollified half-space assembler: edge integrals, moment
// tables, full-space U/H/E kernels.  FP64 throughout; every compile-time loop
// is fully unrolled so the tables stay in registers.
//
// Reference algorithm: src/mhs/fullspace/{primitives,moments,kernels}.py.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MHS_INLINE __host__ __device__ __forceinline__
// The edge integrals and the full direct lift are real device calls, not
// inlined: with them and the near-field moments all inlined into one pair
// kernel, hipcc (ROCm 7.2, -O2/-O3) produced lane-dependent wrong results
// (NaN in one lane whenever the lanes of a wave took different near/far/fold
// branches), while -O1, single-lane launches and any one of the three as a
// call were correct.  The near-field moments are inlined (their 216-entry
// table would otherwise live in scratch memory).
#define MHS_CALL __host__ __device__ __attribute__((noinline))
#ifndef MHS_EDGE_FN
#define MHS_EDGE_FN MHS_CALL
#endif
#ifndef MHS_NEAR_FN
#define MHS_NEAR_FN MHS_INLINE
#endif
#ifndef MHS_DIRECT_FN
#define MHS_DIRECT_FN MHS_CALL
#endif

// ---------------------------------------------------------------- records --
// Per-source record, MHS_SRC_STRIDE doubles (built on the host from the
// vertices alone; the same frame construction as fullspace/frame.py:67-87).
#define MHS_SRC_STRIDE 80
enum SrcOff {
  S_V = 0,      // vertices (3 x 3)
  S_EPS = 9,
  S_E1 = 10, S_E2 = 13, S_N = 16, S_C = 19,   // centroid frame
  S_P = 22,     // in-plane vertex coords about the centroid (3 x 2)
  S_AREA = 28, S_L = 29,
  S_IE1 = 30, S_IE2 = 33, S_IN = 36, S_IC = 39,   // frame of the reflected triangle
  S_ABC = 42,   // barycentric affine rows (A, B, C) per vertex (3 x 3)
  S_CREASE = 51, S_CREASE_R = 53, S_DIAM = 54, S_STRADDLE = 55,
  S_IP = 56,    // reflected triangle in-plane coords (3 x 2)
  S_IABC = 62,  // reflected triangle affine rows
  S_IAREA = 71, S_IL = 72,
  S_RC = 73     // largest vertex distance from the centroid (bounding sphere)
};

// Frame in registers.
struct Frame {
  double e1[3], e2[3], n[3], c[3];
  double p[6];
  double abc[9];
  double area, L;
};

MHS_INLINE void load_frame(const double* __restrict__ s, bool image, Frame& f) {
  const int o1 = image ? S_IE1 : S_E1, o2 = image ? S_IE2 : S_E2;
  const int on = image ? S_IN : S_N, oc = image ? S_IC : S_C;
  const int op = image ? S_IP : S_P, oa = image ? S_IABC : S_ABC;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    f.e1[i] = s[o1 + i]; f.e2[i] = s[o2 + i]; f.n[i] = s[on + i]; f.c[i] = s[oc + i];
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) f.p[i] = s[op + i];
#pragma unroll
  for (int i = 0; i < 9; ++i) f.abc[i] = s[oa + i];
  f.area = image ? s[S_IAREA] : s[S_AREA];
  f.L = image ? s[S_IL] : s[S_L];
}

// 1/sqrt(x) and 1/x in FP64 for the quadrature loops (x > 0, finite): the
// hardware approximations (~2^-22 relative) refined by two Newton steps,
// each squaring the error, to the last bits of a double -- about half the
// instructions of a correctly rounded sqrt and division.
__device__ __forceinline__ double rsq_refined(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double r = fma(-x * y, y, 1.0);          // 1 - x y^2
    y = fma(0.5 * y, r, y);
  }
  return y;
}
__device__ __forceinline__ double rcp_refined(double x) {
  double y = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int it = 0; it < 2; ++it) y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

// Triangular index of xi1^u xi2^v, and table sizes.
__host__ __device__ constexpr int tix(int u, int v) { return (u + v) * (u + v + 1) / 2 + v; }
__host__ __device__ constexpr int tsize(int d) { return d < 0 ? 0 : (d + 1) * (d + 2) / 2; }

// C(-m/2, j) for odd m in [-3, 9] (index (m + 3) / 2) and j < 32: the
// generalised binomial of primitives.py:65-84, built in double at compile
// time by the same product.
struct BinomHalf {
  double v[7][32];
  constexpr BinomHalf() : v() {
    for (int mi = 0; mi < 7; ++mi) {
      const double a = -0.5 * (2 * mi - 3);
      double acc = 1.0;
      for (int j = 0; j < 32; ++j) {
        v[mi][j] = acc;
        acc = acc * (a - j) / (j + 1);
      }
    }
  }
};
static constexpr BinomHalf kBinom{};
#define BINOM(m, j) (kBinom.v[((m) + 3) / 2][(j)])

// ------------------------------------------------------------ degree specs --
// n index (n + 1) / 2 for n = -1, 1, 3, 5, 7, 9; -1 = absent.
// DEG: closed table degree per n (moments.py:94-116 applied to the kernel's
// need at order P); NEED: constant-slip degree of the weighted tables;
// EK: edge-integral kmax per m = n - 2 (moments.py:221-233).
template <int P> struct SpecH {      // direct stress (H) + eigen (E)
  static constexpr int DEG[6] = {-1, P == 2 ? 0 : -1, P, 2 + P, 4 + P, -1};
  static constexpr int NEED[6] = {-1, -1, 0, 2, 4, -1};
  static constexpr int EK[6] = {-1, P >= 1 ? P - 1 : -1, 1 + P, 3 + P, -1, -1};
  static constexpr int KMAX = 3 + P;
};
template <int P> struct SpecU {      // direct displacement (U)
  static constexpr int DEG[6] = {-1, P == 2 ? 1 : (P == 1 ? 0 : -1), 1 + P, 3 + P, -1, -1};
  static constexpr int NEED[6] = {-1, -1, 1, 3, -1, -1};
  static constexpr int EK[6] = {P == 2 ? 0 : -1, P, 2 + P, -1, -1, -1};
  static constexpr int KMAX = 2 + P;
};
template <int P> struct SpecI {      // image R-family: DEGREES {3:1, 5:3, 7:4, 9:5}
  static constexpr int DEG[6] = {-1, P == 2 ? 1 : (P == 1 ? 0 : -1), 1 + P, 3 + P, 4 + P, 5 + P};
  static constexpr int NEED[6] = {-1, -1, 1, 3, 4, 5};
  static constexpr int EK[6] = {P == 2 ? 0 : -1, P, 2 + P, 3 + P, 4 + P, -1};
  static constexpr int KMAX = 4 + P;
};

// ------------------------------------------------------- edge integrals --
// P_k^m = int_{ua}^{ub} u^k (u^2 + rho^2)^(-m/2) du for every (m, k) of the
// spec, in the regime the reference selects per edge (primitives.py:431-504):
//   small |u| series  : max|u| <= rho / 2
//   large |u| series  : same sign and min|u| >= 2 rho
//   closed form       : otherwise.
// The series' power differences are formed without cancellation by the
// factorisation b^n - a^n = (b - a) sum a^i b^(n-1-i) (all terms one sign);
// that equals the reference's expm1/direct forms to rounding.
// Output E[mi][k], mi = (m + 1) / 2 over m = -1 .. 7.
template <class S>
MHS_EDGE_FN void edge_integrals(double ua, double ub, double rho2, double E[5][8]) {
  constexpr int KM = S::KMAX;
  const double rho = sqrt(rho2);
  const double aua = fabs(ua), aub = fabs(ub);
  const double umax = fmax(aua, aub), umin = fmin(aua, aub);
  const bool same = (ua * ub) > 0.0;
  if (umax <= 0.5 * rho) {
    // small-|u| series: rho^(k+1-m) sum_j C(-m/2, j) [y^pw / pw], pw = k + 2j + 1
    const double ya = ua / rho, yb = ub / rho;
    const double dy = yb - ya;
    double acc[5][8];
#pragma unroll
    for (int mi = 0; mi < 5; ++mi)
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[mi][k] = 0.0;
    double S_ = 1.0, yan = ya, ybn = yb;      // S_pw, ya^pw, yb^pw at pw = 1
#pragma unroll
    for (int pw = 1; pw <= KM + 63; ++pw) {
      const double d = same ? dy * S_ : (ybn - yan);
      const double t = d * (1.0 / pw);
#pragma unroll
      for (int mi = 0; mi < 5; ++mi) {
        if (S::EK[mi] < 0) continue;
        const int m = 2 * mi - 1;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (k > S::EK[mi]) continue;
          if (((pw - 1 - k) & 1) || pw - 1 - k < 0) continue;
          const int j = (pw - 1 - k) / 2;
          if (j >= 32) continue;
          acc[mi][k] = fma(BINOM(m, j), t, acc[mi][k]);
        }
      }
      S_ = fma(yb, S_, yan);
      yan *= ya;
      ybn *= yb;
    }
#pragma unroll
    for (int mi = 0; mi < 5; ++mi) {
      if (S::EK[mi] < 0) continue;
      const int m = 2 * mi - 1;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k > S::EK[mi]) continue;
        const int e = k + 1 - m;
        double sc = 1.0;
        if (e > 0) {
#pragma unroll
          for (int i = 0; i < e; ++i) sc *= rho;
        } else {
          const double ir = 1.0 / rho;
#pragma unroll
          for (int i = 0; i < -e; ++i) sc *= ir;
        }
        E[mi][k] = sc * acc[mi][k];
      }
    }
  } else if (same && umin >= 2.0 * rho) {
    // large-|u| series on the positive edge (a, b) = (|.|min, |.|max):
    // a^s sum_j C(-m/2, j) x^j [t^q / q]_1^{b/a}, q = s - 2j, s = k - m + 1.
    const bool neg = ua < 0.0;
    const double a = neg ? -ub : ua, b = neg ? -ua : ub;
    const double x = rho2 / (a * a);
    const double r = b / a, rr = a / b;
    const double dba = b - a;
    const double rm1 = dba / a;            // r - 1
    const double om = dba / b;             // 1 - 1/r
    const double logr = log1p(rm1);
    double xp[32];
    xp[0] = 1.0;
#pragma unroll
    for (int j = 1; j < 32; ++j) xp[j] = xp[j - 1] * x;
    double acc[5][8];
#pragma unroll
    for (int mi = 0; mi < 5; ++mi)
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[mi][k] = 0.0;
    // q from +QMAX down to QMIN; dq(q) = (r^q - 1)/q, log r at q = 0.
    constexpr int QMAX = KM + 2;           // s = k - m + 1 <= KM + 2 (m >= -1)
    constexpr int QMIN = -8 - 62;
    double gp = 0.0, rp = 1.0;             // G+_q = sum_{i<q} r^i
    double dqpos[QMAX + 1];
#pragma unroll
    for (int q = 1; q <= QMAX; ++q) {
      gp = gp + rp;
      rp *= r;
      dqpos[q] = rm1 * gp * (1.0 / q);
    }
    double gm = 0.0, rhn = 1.0;            // G-_n = sum_{i<n} (1/r)^i
#pragma unroll
    for (int q = QMAX; q >= QMIN; --q) {
      double dq;
      if (q > 0) {
        dq = dqpos[q];
      } else if (q == 0) {
        dq = logr;
      } else {
        gm = gm + rhn;
        rhn *= rr;
        dq = om * gm * (1.0 / (-q));
      }
#pragma unroll
      for (int mi = 0; mi < 5; ++mi) {
        if (S::EK[mi] < 0) continue;
        const int m = 2 * mi - 1;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (k > S::EK[mi]) continue;
          const int s = k - m + 1;
          if (((s - q) & 1) || s - q < 0) continue;
          const int j = (s - q) / 2;
          if (j >= 32) continue;
          acc[mi][k] = fma(BINOM(m, j) * dq, xp[j], acc[mi][k]);
        }
      }
    }
#pragma unroll
    for (int mi = 0; mi < 5; ++mi) {
      if (S::EK[mi] < 0) continue;
      const int m = 2 * mi - 1;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k > S::EK[mi]) continue;
        const int s = k - m + 1;
        double base = 1.0;
        if (s > 0) {
#pragma unroll
          for (int i = 0; i < s; ++i) base *= a;
        } else {
          const double ia = 1.0 / a;
#pragma unroll
          for (int i = 0; i < -s; ++i) base *= ia;
        }
        const double v = base * acc[mi][k];
        E[mi][k] = (neg && (k & 1)) ? -v : v;
      }
    }
  } else {
  // closed form (primitives.py:131-212): J_m, K_m and P_k^m = P_{k-2}^{m-2} - rho^2 P_{k-2}^m
  // over m = -3 .. 7 (index (m + 3) / 2).
  const double Ra = sqrt(ua * ua + rho2), Rb = sqrt(ub * ub + rho2);
  const double dR = (ub - ua) * (ub + ua) / (Ra + Rb);
  double J[6], K[6];
  {
    double j1;
    if (ua >= 0.0) j1 = log1p(((ub - ua) + dR) / (ua + Ra));
    else if (ub <= 0.0) j1 = log1p(((ub - ua) - dR) / (Rb - ub));
    else j1 = log(((ub + Rb) * (Ra - ua)) / rho2);
    J[2] = j1;
    double j3;
    if (same) j3 = ((ub - ua) * (ub + ua)) / (Ra * Rb * (ub * Ra + ua * Rb));
    else j3 = (ub / Rb - ua / Ra) / rho2;
    J[3] = j3;
    const double iRa = 1.0 / Ra, iRb = 1.0 / Rb;
    // m >= 5: (u R^-mm |_a^b + (mm - 1) J_mm) / (mm rho2), mm = m - 2
    {
      const double a3 = iRa * iRa * iRa, b3 = iRb * iRb * iRb;
      J[4] = (ub * b3 - ua * a3 + 2.0 * J[3]) / (3.0 * rho2);
      const double a5 = a3 * iRa * iRa, b5 = b3 * iRb * iRb;
      J[5] = (ub * b5 - ua * a5 + 4.0 * J[4]) / (5.0 * rho2);
    }
    // m = -1: (duR + rho2 J_1) / 2;  m = -3: duR^3/4 + 3 rho2/4 J_-1
    const double duR = 0.5 * (ub - ua) * (Ra + Rb) + 0.5 * (ua + ub) * dR;
    J[1] = 0.5 * (duR + rho2 * J[2]);
    J[0] = (ub * Rb * Rb * Rb - ua * Ra * Ra * Ra) / 4.0 + 3.0 * rho2 / 4.0 * J[1];
    // K_m = (R_b^q - R_a^q)/q, q = 2 - m, factorised through dR
#pragma unroll
    for (int mi = 0; mi < 6; ++mi) {
      const int m = 2 * mi - 3;
      const int q = 2 - m;
      if (q > 0) {
        double Ssum = 0.0, ap = 1.0;
#pragma unroll
        for (int i = 0; i < q; ++i) {
          double bp = 1.0;
#pragma unroll
          for (int t = 0; t < q - 1 - i; ++t) bp *= Rb;
          Ssum += ap * bp;
          ap *= Ra;
        }
        K[mi] = dR * Ssum / q;
      } else {
        const int p = -q;
        double Ssum = 0.0, ap = 1.0;
#pragma unroll
        for (int i = 0; i < p; ++i) {
          double bp = 1.0;
#pragma unroll
          for (int t = 0; t < p - 1 - i; ++t) bp *= Rb;
          Ssum += ap * bp;
          ap *= Ra;
        }
        double rab = 1.0;
#pragma unroll
        for (int i = 0; i < p; ++i) rab *= Ra * Rb;
        K[mi] = dR * Ssum / (p * rab);
      }
    }
  }
  double T[6][8];    // T[(m+3)/2][k]
#pragma unroll
  for (int mi = 0; mi < 6; ++mi) { T[mi][0] = J[mi]; T[mi][1] = K[mi]; }
#pragma unroll
  for (int k = 2; k < 8; ++k)
#pragma unroll
    for (int mi = 0; mi < 6; ++mi)
      T[mi][k] = (mi >= 1) ? (T[mi - 1][k - 2] - rho2 * T[mi][k - 2]) : 0.0;
#pragma unroll
  for (int mi = 0; mi < 5; ++mi) {
    if (S::EK[mi] < 0) continue;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (k <= S::EK[mi]) E[mi][k] = T[mi + 1][k];
  }
  }
}

// --------------------------------------------------------- solid angle --
MHS_INLINE double solid_angle(const double* v, const double o[3]) {
  double r1[3], r2[3], r3[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) { r1[i] = v[i] - o[i]; r2[i] = v[3 + i] - o[i]; r3[i] = v[6 + i] - o[i]; }
  const double R1 = sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
  const double R2 = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
  const double R3 = sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2]);
  const double cx = r2[1] * r3[2] - r2[2] * r3[1];
  const double cy = r2[2] * r3[0] - r2[0] * r3[2];
  const double cz = r2[0] * r3[1] - r2[1] * r3[0];
  const double num = r1[0] * cx + r1[1] * cy + r1[2] * cz;
  const double d12 = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
  const double d23 = r2[0] * r3[0] + r2[1] * r3[1] + r2[2] * r3[2];
  const double d13 = r1[0] * r3[0] + r1[1] * r3[1] + r1[2] * r3[2];
  const double den = R1 * R2 * R3 + R3 * d12 + R1 * d23 + R2 * d13;
  return 2.0 * atan2(num, den);
}

// ------------------------------------------------------ raw moment table --
// Closed-form raw moments M_n^(a,b), n in DEG (moments.py:181-355).
// M[ni][tix(a, b)], ni = (n + 1) / 2.
template <class S>
MHS_NEAR_FN void raw_moments_near(const Frame& f, const double* __restrict__ v,
                                 double X1, double X2, double h2, double M[6][36]) {
  constexpr int KM = S::KMAX;
  const double h = sqrt(h2);
  // solid-angle seed at the lifted observer (moments.py:212-219)
  double oe[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) oe[i] = f.c[i] + X1 * f.e1[i] + X2 * f.e2[i] + h * f.n[i];
  const double I3 = -solid_angle(v, oe) / h;
  double BD[5];
  double BN1[5][36];
  double BN2[5][8];
#pragma unroll
  for (int mi = 0; mi < 5; ++mi) {
    BD[mi] = 0.0;
#pragma unroll
    for (int t = 0; t < 36; ++t) BN1[mi][t] = 0.0;
#pragma unroll
    for (int t = 0; t < 8; ++t) BN2[mi][t] = 0.0;
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int e2i = (e + 1) % 3;
    const double pax = f.p[2 * e] - X1, pay = f.p[2 * e + 1] - X2;
    const double pbx = f.p[2 * e2i] - X1, pby = f.p[2 * e2i + 1] - X2;
    const double evx = pbx - pax, evy = pby - pay;
    const double len = sqrt(evx * evx + evy * evy);
    const double s1 = evx / len, s2 = evy / len;
    const double c1 = s2, c2 = -s1;
    const double dp = pax * c1 + pay * c2;
    const double ua = pax * s1 + pay * s2;
    const double ub = pbx * s1 + pby * s2;
    double E[5][8];
    edge_integrals<S>(ua, ub, dp * dp + h2, E);
    // poly[a][b][t]: coefficients of (dp c1 + s1 u)^a (dp c2 + s2 u)^b in u
    const double l10 = dp * c1, l20 = dp * c2;
    double pl[36][8];
    pl[0][0] = 1.0;
#pragma unroll
    for (int t = 1; t < 8; ++t) pl[0][t] = 0.0;
#pragma unroll
    for (int deg = 1; deg <= KM; ++deg) {
#pragma unroll
      for (int b = 0; b <= deg; ++b) {
        const int a = deg - b;
        const int me = tix(a, b);
        if (a > 0) {
          const int par = tix(a - 1, b);
#pragma unroll
          for (int t = 0; t < 8; ++t)
            pl[me][t] = l10 * pl[par][t] + (t > 0 ? s1 * pl[par][t - 1] : 0.0);
        } else {
          const int par = tix(0, b - 1);
#pragma unroll
          for (int t = 0; t < 8; ++t)
            pl[me][t] = l20 * pl[par][t] + (t > 0 ? s2 * pl[par][t - 1] : 0.0);
        }
      }
    }
#pragma unroll
    for (int mi = 0; mi < 5; ++mi) {
      if (S::EK[mi] < 0) continue;
      const int km = S::EK[mi];
      BD[mi] = fma(dp, E[mi][0], BD[mi]);
#pragma unroll
      for (int deg = 0; deg <= km; ++deg) {
#pragma unroll
        for (int b = 0; b <= deg; ++b) {
          const int a = deg - b;
          double val = 0.0;
#pragma unroll
          for (int t = 0; t <= deg; ++t) val = fma(pl[tix(a, b)][t], E[mi][t], val);
          BN1[mi][tix(a, b)] = fma(c1, val, BN1[mi][tix(a, b)]);
          if (a == 0) BN2[mi][b] = fma(c2, val, BN2[mi][b]);
        }
      }
    }
  }
  // seeds (moments.py:303-335)
#pragma unroll
  for (int ni = 0; ni < 6; ++ni)
#pragma unroll
    for (int t = 0; t < 36; ++t) M[ni][t] = 0.0;
  M[2][0] = I3;
  // upward: I_{n+2} = (BD_n + (n - 2) I_n) / (n h^2)
#pragma unroll
  for (int ni = 2; ni < 5; ++ni) {
    if (S::DEG[ni + 1] < 0) break;
    const int n = 2 * ni - 1;
    M[ni + 1][0] = (BD[ni] + (n - 2) * M[ni][0]) / (n * h2);
  }
  // downward: I_n = (BD_n - n h^2 I_{n+2}) / (2 - n), n = 1, -1
#pragma unroll
  for (int ni = 1; ni >= 0; --ni) {
    if (S::DEG[ni] < 0) break;
    const int n = 2 * ni - 1;
    M[ni][0] = (BD[ni] - n * h2 * M[ni + 1][0]) / (2.0 - n);
  }
  // horizontal recursion, ascending n (moments.py:337-354)
#pragma unroll
  for (int ni = 0; ni < 6; ++ni) {
    constexpr int dummy = 0; (void)dummy;
    if (S::DEG[ni] <= 0) continue;
    const int n = 2 * ni - 1;
    const int mi = ni - 1;               // m = n - 2
    const double inv = 1.0 / (n - 2);
#pragma unroll
    for (int q = 1; q <= 7; ++q) {
      if (q > S::DEG[ni]) continue;
#pragma unroll
      for (int a = q; a >= 0; --a) {
        const int b = q - a;
        double val;
        if (a >= 1) {
          val = -BN1[mi][tix(a - 1, b)];
          if (a >= 2) val += (a - 1) * M[mi][tix(a - 2, b)];
        } else {
          val = -BN2[mi][b - 1];
          if (b >= 2) val += (b - 1) * M[mi][tix(0, b - 2)];
        }
        M[ni][tix(a, b)] = val * inv;
      }
    }
  }
}

// ------------------------------------------------------ shape functions --
// Lagrange N_k about the observer foot X as monomials in xi (shape.py:238-271):
// c[tix(a, b)] for a + b <= P.
template <int P>
MHS_INLINE void shape_coeffs(const Frame& f, int k, double X1, double X2, double c[6]) {
  if (P == 0) { c[0] = 1.0; return; }
  double l0[3], lx[3], ly[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    lx[i] = f.abc[3 * i + 1];
    ly[i] = f.abc[3 * i + 2];
    l0[i] = f.abc[3 * i] + X1 * lx[i] + X2 * ly[i];
  }
  if (P == 1) {
    c[0] = l0[k]; c[1] = lx[k]; c[2] = ly[k];
    return;
  }
  if (k < 3) {                    // lam (2 lam - 1)
    const double a0 = l0[k], ax = lx[k], ay = ly[k];
    c[0] = a0 * (2.0 * a0 - 1.0);
    c[1] = ax * (4.0 * a0 - 1.0);
    c[2] = ay * (4.0 * a0 - 1.0);
    c[3] = 2.0 * ax * ax;
    c[4] = 4.0 * ax * ay;
    c[5] = 2.0 * ay * ay;
  } else {                        // 4 lam_i lam_j, edges 12, 23, 31
    const int i = k - 3, j = (k - 2) % 3;
    c[0] = 4.0 * l0[i] * l0[j];
    c[1] = 4.0 * (l0[i] * lx[j] + lx[i] * l0[j]);
    c[2] = 4.0 * (l0[i] * ly[j] + ly[i] * l0[j]);
    c[3] = 4.0 * lx[i] * lx[j];
    c[4] = 4.0 * (lx[i] * ly[j] + ly[i] * lx[j]);
    c[5] = 4.0 * ly[i] * ly[j];
  }
}

// N_k at barycentric (l0, l1, l2).
template <int P>
MHS_INLINE double shape_at(int k, double l0, double l1, double l2) {
  if (P == 0) return 1.0;
  const double l[3] = {l0, l1, l2};
  if (P == 1) return l[k];
  if (k < 3) return l[k] * (2.0 * l[k] - 1.0);
  return 4.0 * l[k - 3] * l[(k - 2) % 3];
}

// Weighted table W[ni][tix] of node k from the raw closed-form table
// (moments.py:370-398).
template <class S, int P>
MHS_INLINE void weight_from_raw(const double M[6][36], const double c[6], double W[6][21]) {
#pragma unroll
  for (int ni = 0; ni < 6; ++ni) {
    if (S::NEED[ni] < 0) continue;
#pragma unroll
    for (int d = 0; d <= 5; ++d) {
      if (d > S::NEED[ni]) continue;
#pragma unroll
      for (int b = 0; b <= d; ++b) {
        const int a = d - b;
        double acc = 0.0;
#pragma unroll
        for (int dd = 0; dd <= P; ++dd)
#pragma unroll
          for (int bp = 0; bp <= dd; ++bp) {
            const int ap = dd - bp;
            acc = fma(c[tix(ap, bp)], M[ni][tix(a + ap, b + bp)], acc);
          }
        W[ni][tix(a, b)] = acc;
      }
    }
  }
}

// Far-field weighted table of node k by the collapsed Gauss rule
// (moments.py:427-498): W[n][a, b] = sum_q w_q N_k(q) xi1^a xi2^b R^-n.
template <class S, int P>
MHS_INLINE void weighted_far(const Frame& f, int k, double X1, double X2, double h2, int ng,
                             const double* __restrict__ gl_x, const double* __restrict__ gl_w,
                             double W[6][21]) {
#pragma unroll
  for (int ni = 0; ni < 6; ++ni)
#pragma unroll
    for (int t = 0; t < 21; ++t) W[ni][t] = 0.0;
  const double ar2 = 2.0 * f.area;
  for (int i = 0; i < ng; ++i) {
    const double gi = gl_x[i], wi = gl_w[i];
    for (int j = 0; j < ng; ++j) {
      const double x1 = gi, x2 = gl_x[j] * (1.0 - gi);
      const double w = wi * gl_w[j] * (1.0 - gi) * ar2;
      const double l0 = 1.0 - x1 - x2;
      const double eta1 = l0 * f.p[0] + x1 * f.p[2] + x2 * f.p[4];
      const double eta2 = l0 * f.p[1] + x1 * f.p[3] + x2 * f.p[5];
      const double wn = w * shape_at<P>(k, l0, x1, x2);
      const double xi1 = eta1 - X1, xi2 = eta2 - X2;
      const double R2 = xi1 * xi1 + xi2 * xi2 + h2;
      const double ir = rsq_refined(R2);
      const double ir2 = ir * ir;
      double rn[6];
      rn[0] = R2 * ir;  // R^1 (unused)
      rn[1] = ir;
#pragma unroll
      for (int ni = 2; ni < 6; ++ni) rn[ni] = rn[ni - 1] * ir2;
      double mono[21];
      mono[0] = wn;
#pragma unroll
      for (int d = 1; d <= 5; ++d)
#pragma unroll
        for (int b = 0; b <= d; ++b) {
          const int a = d - b;
          mono[tix(a, b)] = (a > 0) ? mono[tix(a - 1, b)] * xi1 : mono[tix(0, b - 1)] * xi2;
        }
#pragma unroll
      for (int ni = 0; ni < 6; ++ni) {
        if (S::NEED[ni] < 0) continue;
#pragma unroll
        for (int t = 0; t < 21; ++t)
          if (t < tsize(S::NEED[ni])) W[ni][t] = fma(mono[t], rn[ni], W[ni][t]);
      }
    }
  }
}

// sqrt(|o - c|^2 + eps^2), the near/far switch distance (moments.py:551-554),
// with the rounding pinned by explicit FMAs so that every kernel that
// classifies a pair (the near pass, the far pass) takes the same branch.
MHS_INLINE double centroid_distance(const double* c, const double o[3], double eps) {
  const double d0 = o[0] - c[0], d1 = o[1] - c[1], d2 = o[2] - c[2];
  return sqrt(fma(d0, d0, fma(d1, d1, fma(d2, d2, eps * eps))));
}

// Quadrature rule of the mixed mode.  A pair is near (closed form) when
// sqrt(|x - c|^2 + eps^2) <= near L; otherwise each far-field family
// (direct on the source triangle; image R and Q on the mirrored one, from
// its centroid) takes the Gauss order of the first band with that distance
// <= hi[b] L, plus the nodal order P.  The all-FP64 mode keeps the
// reference's rule (near 10 L, 12 x 12 to 40 L, 8 x 8 beyond, the Q budget
// law) and never reads this.
#ifndef MHS_RULE_BANDS
#define MHS_RULE_BANDS 8
#endif
struct Rule {
  double near;
  double hi[MHS_RULE_BANDS];
  int n[MHS_RULE_BANDS];
};

// Gauss order of a family at distance D from a triangle of size L.
MHS_INLINE int rule_order(const Rule& r, double D, double L) {
#pragma unroll
  for (int b = 0; b < MHS_RULE_BANDS - 1; ++b)
    if (D <= r.hi[b] * L) return r.n[b];
  return r.n[MHS_RULE_BANDS - 1];
}

// Weighted tables of node k by the hybrid switch (moments.py:551-601);
// returns the plane height z.  With a mixed rule: its near switch and far
// orders instead of 10 L and 12^2 / 8^2.
// Coordinates of o in the frame f: returns the normal offset z; X1, X2 in
// the plane, h2 = z^2 + eps^2.
MHS_INLINE double frame_local(const Frame& f, const double o[3], double eps, double& X1, double& X2,
                              double& h2) {
  const double d0 = o[0] - f.c[0], d1 = o[1] - f.c[1], d2 = o[2] - f.c[2];
  const double z = d0 * f.n[0] + d1 * f.n[1] + d2 * f.n[2];
  X1 = d0 * f.e1[0] + d1 * f.e1[1] + d2 * f.e1[2];
  X2 = d0 * f.e2[0] + d1 * f.e2[1] + d2 * f.e2[2];
  h2 = z * z + eps * eps;
  return z;
}

template <class S, int P>
MHS_INLINE double weighted_tables(const Frame& f, const double* __restrict__ v, int k,
                                  const double o[3], double eps,
                                  const double* __restrict__ gl_x, const double* __restrict__ gl_w,
                                  double W[6][21], double& X1o, double& X2o, double near_r = 10.0,
                                  const Rule* rule = nullptr) {
  double X1, X2, h2;
  const double z = frame_local(f, o, eps, X1, X2, h2);
  X1o = X1; X2o = X2;
  const double Dn = centroid_distance(f.c, o, eps);
  if (Dn <= near_r * f.L) {
    double M[6][36];
    raw_moments_near<S>(f, v, X1, X2, h2, M);
    double c[6];
    shape_coeffs<P>(f, k, X1, X2, c);
    weight_from_raw<S, P>(M, c, W);
  } else {
    const int ng = rule ? rule_order(*rule, Dn, f.L) + P : ((Dn > 40.0 * f.L) ? 8 : 12);
    const int off = ng * (ng - 1) / 2;
    weighted_far<S, P>(f, k, X1, X2, h2, ng, gl_x + off, gl_w + off, W);
  }
  return z;
}

// Signed local-frame tensor moment: int N d_loc^(c1, c2, c3) / R^n with
// d_loc = (-xi1, -xi2, z): (-1)^(c1 + c2) z^c3 W[n][c1, c2].
MHS_INLINE double tmom(const double* Wn, int c1, int c2, int c3, const double* zp) {
  const double s = ((c1 + c2) & 1) ? -1.0 : 1.0;
  return s * zp[c3] * Wn[tix(c1, c2)];
}

// Rotate a local (e1, e2, n) rank-1 index to global.
MHS_INLINE void basis_matrix(const Frame& f, double R[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) { R[i][0] = f.e1[i]; R[i][1] = f.e2[i]; R[i][2] = f.n[i]; }
}

struct Mat {
  double mu, lam, nu;
};

// Full-space TOTAL stress H[m][l][j] (global) and blob weight E of node k
// (kernels.py:182-213), evaluated in the source's local frame where the
// normal is e3, then rotated.
MHS_DIRECT_FN void direct_H(const Frame& f, const double W[6][21], double z, double eps, const Mat& mt,
                         double H[27]) {
  const double nu = mt.nu, mu = mt.mu, lam = 2.0 * mu * nu / (1.0 - 2.0 * nu);
  const double C1 = 1.0 / (16.0 * 3.14159265358979323846 * mu * (1.0 - nu));
  const double c34 = 3.0 - 4.0 * nu;
  const double cb = 2.0 * (1.0 - nu) * eps * eps;
  double zp[5] = {1.0, z, z * z, z * z * z, z * z * z * z};
  const double I3 = W[2][0], I5 = W[3][0];
  // local tensors by index counts
  auto cnt = [](int a, int b, int c, int d, int which) {
    return (a == which) + (b == which) + (c == which) + (d == which);
  };
  double T25[3][3], T27[3][3];
#pragma unroll
  for (int s = 0; s < 3; ++s)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int c1 = (s == 0) + (q == 0), c2 = (s == 1) + (q == 1), c3 = (s == 2) + (q == 2);
      T25[s][q] = tmom(W[3], c1, c2, c3, zp);
      T27[s][q] = tmom(W[4], c1, c2, c3, zp);
    }
  // B[r][s][j] = lam d_j3 trD[r][s] + mu D2[r][j][s][3] + mu D2[r][3][s][j]
  // with D2 the local second-derivative tensor (kernels.py:188-202).
  auto D2 = [&](int r, int p, int s, int q) -> double {
    const double drp = r == p, drs = r == s, dps = p == s, drq = r == q, dpq = p == q, dsq = s == q;
    double v = -c34 * drp * (dsq * I3 - 3.0 * T25[s][q]);
    v += drs * (dpq * I3 - 3.0 * T25[p][q]);
    v += dps * (drq * I3 - 3.0 * T25[r][q]);
    v -= 3.0 * (drq * T25[p][s] + dpq * T25[r][s] + dsq * T25[r][p]);
    v += 15.0 * tmom(W[4], cnt(r, p, s, q, 0), cnt(r, p, s, q, 1), cnt(r, p, s, q, 2), zp);
    v += cb * drp * (-3.0 * dsq * I5 + 15.0 * T27[s][q]);
    return C1 * v;
  };
  double B[3][3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const double trD = D2(r, 0, s, 0) + D2(r, 1, s, 1) + D2(r, 2, s, 2);
#pragma unroll
      for (int j = 0; j < 3; ++j)
        B[r][s][j] = (j == 2 ? lam * trD : 0.0) + mu * D2(r, j, s, 2) + mu * D2(r, 2, s, j);
    }
  double Hl[3][3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double trB = B[0][0][j] + B[1][1][j] + B[2][2][j];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int l = 0; l < 3; ++l)
        Hl[m][l][j] = -((m == l ? lam * trB : 0.0) + mu * (B[m][l][j] + B[l][m][j]));
  }
  double R[3][3];
  basis_matrix(f, R);
  // H_glob[M][L][J] = R[M][m] R[L][l] R[J][j] Hl[m][l][j]
  double A1[3][3][3], A2[3][3][3];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int J = 0; J < 3; ++J)
        A1[m][l][J] = R[J][0] * Hl[m][l][0] + R[J][1] * Hl[m][l][1] + R[J][2] * Hl[m][l][2];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int L = 0; L < 3; ++L)
#pragma unroll
      for (int J = 0; J < 3; ++J)
        A2[m][L][J] = R[L][0] * A1[m][0][J] + R[L][1] * A1[m][1][J] + R[L][2] * A1[m][2][J];
#pragma unroll
  for (int M_ = 0; M_ < 3; ++M_)
#pragma unroll
    for (int L = 0; L < 3; ++L)
#pragma unroll
      for (int J = 0; J < 3; ++J)
        H[(M_ * 3 + L) * 3 + J] = R[M_][0] * A2[0][L][J] + R[M_][1] * A2[1][L][J] + R[M_][2] * A2[2][L][J];
}

// Receiver and source halves of a contracted output, symmetric 3 x 3 in the
// order 00 11 22 01 02 12 (global frame):
//   E = mu (rb nr^T + nr rb^T) + lam (rb . nr) I   (traction along rb on the
//       plane of normal nr, as a functional of strain);
//   F = mu (sb n^T + n sb^T) + lam (sb . n) I      (Hooke of slip sb on the
//       source plane of normal n).
MHS_INLINE void half_tensor(const double* u, const double* n, double mu, double lam, double A[6]) {
  const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
  A[0] = 2.0 * mu * u[0] * n[0] + lam * un;
  A[1] = 2.0 * mu * u[1] * n[1] + lam * un;
  A[2] = 2.0 * mu * u[2] * n[2] + lam * un;
  A[3] = mu * (u[0] * n[1] + n[0] * u[1]);
  A[4] = mu * (u[0] * n[2] + n[0] * u[2]);
  A[5] = mu * (u[1] * n[2] + n[1] * u[2]);
}

// d.A.d for the local offset d = (-xi1, -xi2, z) of the frame f as a
// polynomial in (xi1, xi2), coefficients in table order tix(u, v).
MHS_INLINE void local_quadratic(const Frame& f, const double A[6], double z, double p[6]) {
  const double M[3][3] = {{A[0], A[3], A[4]}, {A[3], A[1], A[5]}, {A[4], A[5], A[2]}};
  const double* Q[3] = {f.e1, f.e2, f.n};
  double L[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a)
        s += Q[i][a] * (M[a][0] * Q[j][0] + M[a][1] * Q[j][1] + M[a][2] * Q[j][2]);
      L[i][j] = s;
    }
  p[tix(0, 0)] = z * z * L[2][2];
  p[tix(1, 0)] = -2.0 * z * L[0][2];
  p[tix(0, 1)] = -2.0 * z * L[1][2];
  p[tix(2, 0)] = L[0][0];
  p[tix(1, 1)] = 2.0 * L[0][1];
  p[tix(0, 2)] = L[1][1];
}

// The direct (Kelvin-Cortez) total stress of node k contracted with a
// receiver half E and a source half F -- rb.H.nr for slip sb -- from the
// weighted tables: the pointwise kernel
//   -C1 [A3/R^3 + (A5c c + A5a a + A5b b + A5e)/R^5 + (15 a b + A7c c)/R^7]
// with a = d.E.d, b = d.F.d, c = d.sym(EF).d (the contraction of
// kernels.py:182-213's H with rb, nr and sb), each a quadratic in the in-plane
// offset, integrated against W5 (degree <= 2) and W7 (degree <= 4).  With
// eig, minus the eigenstress term E_blob * rb.F.nr (assemble.py:123-149).
MHS_INLINE double direct_contract(const Frame& f, const double W[6][21], double z, double eps,
                                  const Mat& mt, const double E[6], const double F[6], bool eig,
                                  double rFn) {
  const double nu = mt.nu;
  const double C1 = 1.0 / (16.0 * 3.14159265358979323846 * mt.mu * (1.0 - nu));
  const double c34 = 3.0 - 4.0 * nu;
  const double Em[3][3] = {{E[0], E[3], E[4]}, {E[3], E[1], E[5]}, {E[4], E[5], E[2]}};
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
  const double trE = E[0] + E[1] + E[2], trF = F[0] + F[1] + F[2];
  double pa[6], pb[6], pc[6];
  local_quadratic(f, E, z, pa);
  local_quadratic(f, F, z, pb);
  local_quadratic(f, Ms, z, pc);
  const double e2 = eps * eps;
  const double A3 = (1.0 - c34) * EF + trE * trF;
  const double A5c = 3.0 * c34 - 9.0, A5a = -3.0 * trF, A5b = -3.0 * trE;
  const double A5e = -6.0 * (1.0 - nu) * e2 * EF;
  const double A7c = 30.0 * (1.0 - nu) * e2;
  const double* W3 = W[2];
  const double* W5 = W[3];
  const double* W7 = W[4];
  double S = A3 * W3[0] + A5e * W5[0];
  constexpr int U[6] = {0, 1, 0, 2, 1, 0}, V[6] = {0, 0, 1, 0, 1, 2};   // exponents of tix 0..5
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    S = fma(A5c * pc[t] + A5a * pa[t] + A5b * pb[t], W5[t], S);
    S = fma(A7c * pc[t], W7[t], S);
  }
#pragma unroll
  for (int t1 = 0; t1 < 6; ++t1)
#pragma unroll
    for (int t2 = 0; t2 < 6; ++t2)
      S = fma(15.0 * pa[t1] * pb[t2], W7[tix(U[t1] + U[t2], V[t1] + V[t2])], S);
  double out = -C1 * S;
  if (eig) out -= (15.0 * e2 * e2 / (8.0 * 3.14159265358979323846)) * W7[0] * rFn;
  return out;
}

MHS_INLINE double direct_E(const double W[6][21], double eps) {
  return (15.0 * eps * eps * eps * eps / (8.0 * 3.14159265358979323846)) * W[4][0];
}

// Full-space displacement U[i][j] (global) of node k (kernels.py:148-163).
MHS_DIRECT_FN void direct_U(const Frame& f, const double W[6][21], double z, double eps, const Mat& mt,
                         double U[9]) {
  const double nu = mt.nu, mu = mt.mu, lam = 2.0 * mu * nu / (1.0 - 2.0 * nu);
  const double C1 = 1.0 / (16.0 * 3.14159265358979323846 * mu * (1.0 - nu));
  const double c34 = 3.0 - 4.0 * nu;
  const double cb6 = 6.0 * (1.0 - nu) * eps * eps;
  double zp[5] = {1.0, z, z * z, z * z * z, z * z * z * z};
  double V3[3], V5[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    V3[m] = tmom(W[2], m == 0, m == 1, m == 2, zp);
    V5[m] = tmom(W[3], m == 0, m == 1, m == 2, zp);
  }
  auto G1 = [&](int i, int j, int m) -> double {
    const double dij = i == j, dim = i == m, djm = j == m;
    const int c1 = (i == 0) + (j == 0) + (m == 0), c2 = (i == 1) + (j == 1) + (m == 1);
    const int c3 = 3 - c1 - c2;
    double v = -c34 * dij * V3[m] + dim * V3[j] + djm * V3[i] - 3.0 * tmom(W[3], c1, c2, c3, zp)
               - cb6 * dij * V5[m];
    return C1 * v;
  };
  double Ul[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double tr = G1(i, 0, 0) + G1(i, 1, 1) + G1(i, 2, 2);
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Ul[i][j] = -(mu * G1(i, j, 2) + (j == 2 ? lam * tr : 0.0) + mu * G1(i, 2, j));
  }
  double R[3][3];
  basis_matrix(f, R);
  double A[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int J = 0; J < 3; ++J) A[i][J] = R[J][0] * Ul[i][0] + R[J][1] * Ul[i][1] + R[J][2] * Ul[i][2];
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int J = 0; J < 3; ++J) U[I * 3 + J] = R[I][0] * A[0][J] + R[I][1] * A[1][J] + R[I][2] * A[2][J];
}

// Image-family contraction (image.py:107-162): DDG[i][j][p][m] -> H[a][b][k];
// DG[i][j][m] -> U[i][k]; nrm is the real element's vertex-order normal.
// Hooke on the receiver side of the traction-operator tensor S[i][p][k]:
// H[a][b][k] = 2 mu sym(S)[a][b] + lam tr(sym S) delta_ab, per slip k.
MHS_INLINE void stress_from_S(const double* S, double mu, double lam, double H[27]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double e[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) e[a][b] = 0.5 * (S[(a * 3 + b) * 3 + k] + S[(b * 3 + a) * 3 + k]);
    const double tr = e[0][0] + e[1][1] + e[2][2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) H[(a * 3 + b) * 3 + k] = 2.0 * mu * e[a][b] + (a == b ? lam * tr : 0.0);
  }
}

MHS_INLINE void contract_H(const double* DDG, const double* nrm, double mu, double lam, double H[27]) {
  double S[3][3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      double trm = 0.0;
#pragma unroll
      for (int m = 0; m < 3; ++m) trm += DDG[((i * 3 + m) * 3 + p) * 3 + m];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          a += nrm[l] * DDG[((i * 3 + k) * 3 + p) * 3 + l];
          b += nrm[l] * DDG[((i * 3 + l) * 3 + p) * 3 + k];
        }
        S[i][p][k] = mu * a + mu * b + lam * nrm[k] * trm;
      }
    }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double e[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) e[a][b] = 0.5 * (S[a][b][k] + S[b][a][k]);
    const double tr = e[0][0] + e[1][1] + e[2][2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) H[(a * 3 + b) * 3 + k] = 2.0 * mu * e[a][b] + (a == b ? lam * tr : 0.0);
  }
}

MHS_INLINE void contract_U(const double* DG, const double* nrm, double mu, double lam, double U[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double tr = DG[(i * 3 + 0) * 3 + 0] + DG[(i * 3 + 1) * 3 + 1] + DG[(i * 3 + 2) * 3 + 2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        a += nrm[l] * DG[(i * 3 + k) * 3 + l];
        b += nrm[l] * DG[(i * 3 + l) * 3 + k];
      }
      U[i * 3 + k] = mu * a + mu * b + lam * nrm[k] * tr;
    }
  }
}
