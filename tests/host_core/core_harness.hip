// Host build of the per-pair device math (csrc/core.hpp, csrc/pair.hpp) for
// tests that run without a GPU: the direct nodal stress and displacement
// (edge integrals, closed-form near moments or Gauss far tables, the lift)
// and the image R-family, for one (observer, source, node) pair at a time,
// exactly the functions the pair kernels call.  Built twice by
// tests/test_core_host.py under the host sanitizers: with the library's
// call boundaries (edge integrals and the direct lift as real calls) and
// with everything inlined.
//
// stdin: int32 P, int32 n_cases; float64 mu, lam; float64 gl_x[18528],
// gl_w[18528] (Gauss-Legendre orders 1..192 on [0, 1], concatenated); then
// per case float64 src[80], obs[3], int32 k.
// stdout: per case float64 H[27], phi, U[9].
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

// Host counterparts of the refined device reciprocals (the device versions
// use the hardware approximations and are not compiled for the host).
__host__ inline double rsq_refined(double x) { return 1.0 / sqrt(x); }
__host__ inline double rcp_refined(double x) { return 1.0 / x; }

#include "../../mhs_amd/csrc/pair.hpp"

static const int kGauss = 192 * 193 / 2;

template <int P>
static void run_case(const Params& pr, const double* src, const double* o, int k, double* out) {
  double H[27], U[9], phi_d = 0.0, phi_m = 0.0;
  pair_stress<P>(pr, src, k, o, false, true, H, phi_d, phi_m);
  pair_disp<P>(pr, src, k, o, false, U);
  memcpy(out, H, sizeof H);
  out[27] = phi_d;
  memcpy(out + 28, U, sizeof U);
}

static void read_exact(void* p, size_t n) {
  if (fread(p, 1, n, stdin) != n) {
    fprintf(stderr, "core_harness: short input\n");
    exit(2);
  }
}

int main() {
  int P = 0, n = 0;
  read_exact(&P, 4);
  read_exact(&n, 4);
  double mat[2];
  read_exact(mat, sizeof mat);
  std::vector<double> gx(kGauss), gw(kGauss), coef(MHS_NCOEF);
  read_exact(gx.data(), 8 * kGauss);
  read_exact(gw.data(), 8 * kGauss);
  Params pr;
  memset(&pr, 0, sizeof pr);
  pr.mu = mat[0];
  pr.lam = mat[1];
  pr.nu = 0.5 * pr.lam / (pr.lam + pr.mu);
  mhs_fill_coeffs(pr.mu, pr.lam, coef.data());
  pr.coef = coef.data();
  pr.gl_x = gx.data();
  pr.gl_w = gw.data();
  pr.near_r = 10.0;
  pr.mixed_rule = 0;
  pr.flags = F_FOLD;
  double src[MHS_SRC_STRIDE], o[3], out[37];
  for (int c = 0; c < n; ++c) {
    int k = 0;
    read_exact(src, sizeof src);
    read_exact(o, sizeof o);
    read_exact(&k, 4);
    if (P == 0) run_case<0>(pr, src, o, k, out);
    else if (P == 1) run_case<1>(pr, src, o, k, out);
    else run_case<2>(pr, src, o, k, out);
    fwrite(out, sizeof out, 1, stdout);
  }
  return 0;
}
