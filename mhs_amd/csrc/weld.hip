// Shared-node table of the continuous layout on the device: the vertex weld,
// the edge table and the per-node sums of src/mhs/shared.py:175-273 for
// large fault networks.
//
// The host restatement keys every corner by (round(x / tol), round(y / tol),
// round(z / tol), group) and takes np.unique of the rows; edges are
// np.unique of the sorted vertex-id pairs.  Here both are least-significant-
// column-first stable radix sorts (rocPRIM) of the flat corner (or edge
// slot) positions, then a scan of "new key" flags: the unique rows come out
// in the same lexicographic order, so every node id equals the host's.  Each
// node's occurrences are then a run of the sorted positions in increasing
// flat order -- the order np.add.at visits them -- so the vertex averages,
// the area-weighted normals and the edge weights are summed in the host's
// order (no contraction into FMAs: plain IEEE adds, multiplies, divisions).
//
// Each column is sorted on its offset from the column's minimum over only
// the bits its span needs (the group column of a one-fault mesh needs none;
// vertex ids need log2(n_v)), instead of all 64.
//
// Memory: the caller passes every output at its capacity (3 n_tri vertices,
// 3 n_tri edges) and one workspace of mhs_weld_workspace_bytes(n_tri) bytes;
// nothing is allocated here.  The call synchronises its stream four times,
// to read the weld keys' ranges, the two node counts and the refusal bits.
#include <cstdint>
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "mhs_amd.h"

#pragma clang fp contract(off)

int mhs_internal_fail(const char* msg);
int mhs_internal_check(hipError_t e, const char* what);

namespace {

constexpr int TB = 256;

unsigned blocks_for(int64_t n) { return (unsigned)((n + TB - 1) / TB); }

// OR of the wave's active lanes' bits, at most one atomic per wave, by its
// first active lane
__device__ void wave_or(unsigned long long* flags, unsigned bits) {
  unsigned long long any = 0;
  for (unsigned b = 1; b <= 8; b <<= 1)
    if (__ballot((bits & b) != 0)) any |= b;
  const unsigned long long active = __ballot(1);
  if (any && (int)(threadIdx.x & 63) == __ffsll((long long)active) - 1 &&
      (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & any) != any)
    atomicOr(flags, any);   // only waves with a bit not yet set
}

// Per triangle, as the host table computes them: the shortest edge (edge k
// = v_k - v_{k-1}), the largest |coordinate|, non-finite and collinear
// counts reduced into stats[0..3] (doubles as bit patterns: for
// non-negative values the unsigned order is the numeric order), and the
// area-weighted upward unit normal and area / 3 (tdcs.py:49-66).
__global__ void k_tri_geometry(const double* __restrict__ tris, int64_t n_tri, double* __restrict__ an,
                               double* __restrict__ area3, unsigned long long* __restrict__ stats) {
  const int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
  double hmin = __longlong_as_double(0x7ff0000000000000LL), amax = 0.0;
  unsigned long long nonfinite = 0, collinear = 0;
  if (t < n_tri) {
    const double* v = tris + 9 * t;
    bool fin = true;
    for (int i = 0; i < 9; ++i) {
      fin = fin && isfinite(v[i]);
      amax = fmax(amax, fabs(v[i]));
    }
    nonfinite = !fin;
    for (int k = 0; k < 3; ++k) {
      const double* a = v + 3 * k;
      const double* b = v + 3 * (k == 0 ? 2 : k - 1);
      const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
      hmin = fmin(hmin, sqrt(dx * dx + dy * dy + dz * dz));
    }
    const double a0 = v[3] - v[0], a1 = v[4] - v[1], a2 = v[5] - v[2];
    const double b0 = v[6] - v[0], b1 = v[7] - v[1], b2 = v[8] - v[2];
    const double c0 = a1 * b2 - a2 * b1, c1 = a2 * b0 - a0 * b2, c2 = a0 * b1 - a1 * b0;
    const double norm = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    collinear = !(norm > 0.0);
    const double area = 0.5 * norm;
    double n0 = c0 / norm, n1 = c1 / norm, n2 = c2 / norm;
    if (n2 < 0.0) { n0 = -n0; n1 = -n1; n2 = -n2; }
    an[3 * t] = area * n0;
    an[3 * t + 1] = area * n1;
    an[3 * t + 2] = area * n2;
    area3[t] = area / 3.0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    hmin = fmin(hmin, __shfl_down(hmin, off, 64));
    amax = fmax(amax, __shfl_down(amax, off, 64));
    nonfinite += __shfl_down(nonfinite, off, 64);
    collinear += __shfl_down(collinear, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    // only waves that improve on the value seen so far touch the counters
    const unsigned long long hb = (unsigned long long)__double_as_longlong(hmin);
    const unsigned long long ab = (unsigned long long)__double_as_longlong(amax);
    if (hb < __hip_atomic_load(&stats[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&stats[0], hb);
    if (ab > __hip_atomic_load(&stats[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&stats[1], ab);
    if (nonfinite) atomicAdd(&stats[2], nonfinite);
    if (collinear) atomicAdd(&stats[3], collinear);
  }
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_stats_init(unsigned long long* __restrict__ stats) {
  if (threadIdx.x == 0) {
    stats[0] = 0x7ff0000000000000ULL;   // +inf
    stats[1] = 0;
    stats[2] = 0;
    stats[3] = 0;
  }
}

// Order-preserving map of int64 to uint64 (the sign bit flipped), for the
// key ranges' unsigned atomics.
__device__ unsigned long long ordered(int64_t x) { return (unsigned long long)x ^ 0x8000000000000000ULL; }

__global__ void k_weld_keys(const double* __restrict__ tris, const int64_t* __restrict__ groups, int64_t n3,
                            double tol, int64_t* __restrict__ kx, int64_t* __restrict__ ky,
                            int64_t* __restrict__ kz, int64_t* __restrict__ kg, int64_t* __restrict__ pos,
                            unsigned long long* __restrict__ partial) {
  const int64_t j = (int64_t)blockIdx.x * TB + threadIdx.x;
  unsigned long long lo[4] = {~0ULL, ~0ULL, ~0ULL, ~0ULL}, hi[4] = {0, 0, 0, 0};
  if (j < n3) {
    // np.round (half to even) of the IEEE quotient, then the exact conversion
    const int64_t k[4] = {(int64_t)rint(tris[3 * j] / tol), (int64_t)rint(tris[3 * j + 1] / tol),
                          (int64_t)rint(tris[3 * j + 2] / tol), groups[j / 3]};
    kx[j] = k[0];
    ky[j] = k[1];
    kz[j] = k[2];
    kg[j] = k[3];
    pos[j] = j;
    for (int c = 0; c < 4; ++c) lo[c] = hi[c] = ordered(k[c]);
  }
  // each column's range: wave, then workgroup minimum / maximum into this
  // workgroup's partial (k_range_reduce takes the whole range from those)
  __shared__ unsigned long long red[TB / 64][8];
  for (int c = 0; c < 4; ++c)
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long a = __shfl_down(lo[c], off, 64), b = __shfl_down(hi[c], off, 64);
      lo[c] = a < lo[c] ? a : lo[c];
      hi[c] = b > hi[c] ? b : hi[c];
    }
  if ((threadIdx.x & 63) == 0)
    for (int c = 0; c < 4; ++c) {
      red[threadIdx.x >> 6][2 * c] = lo[c];
      red[threadIdx.x >> 6][2 * c + 1] = hi[c];
    }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int q = threadIdx.x;
    unsigned long long v = red[0][q];
    for (int wv = 1; wv < TB / 64; ++wv)
      v = (q & 1) ? (red[wv][q] > v ? red[wv][q] : v) : (red[wv][q] < v ? red[wv][q] : v);
    partial[8 * (int64_t)blockIdx.x + q] = v;
  }
}

// The weld keys' ranges from the workgroups' partials: range[2c] = minimum,
// range[2c + 1] = maximum of column c (order-mapped).  One workgroup.
__global__ void k_range_reduce(const unsigned long long* __restrict__ partial, int64_t n_blocks,
                               unsigned long long* __restrict__ range) {
  __shared__ unsigned long long red[TB / 64][8];
  unsigned long long v[8];
  for (int q = 0; q < 8; ++q) v[q] = (q & 1) ? 0ULL : ~0ULL;
  for (int64_t b = threadIdx.x; b < n_blocks; b += TB)
    for (int q = 0; q < 8; ++q) {
      const unsigned long long x = partial[8 * b + q];
      v[q] = (q & 1) ? (x > v[q] ? x : v[q]) : (x < v[q] ? x : v[q]);
    }
  for (int q = 0; q < 8; ++q)
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long x = __shfl_down(v[q], off, 64);
      v[q] = (q & 1) ? (x > v[q] ? x : v[q]) : (x < v[q] ? x : v[q]);
    }
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 8; ++q) red[threadIdx.x >> 6][q] = v[q];
  __syncthreads();
  if (threadIdx.x < 8) {
    const int q = threadIdx.x;
    unsigned long long x = red[0][q];
    for (int wv = 1; wv < TB / 64; ++wv)
      x = (q & 1) ? (red[wv][q] > x ? red[wv][q] : x) : (red[wv][q] < x ? red[wv][q] : x);
    range[q] = x;
  }
}

// local edge slots (v0 v1) (v1 v2) (v2 v0) per triangle, as (min, max); the
// welded vertex ids are columns 0..2 of the (n_tri, 6) id table
__global__ void k_edge_keys(const int64_t* __restrict__ ids, int64_t n3, int64_t* __restrict__ ka,
                            int64_t* __restrict__ kb, int* __restrict__ fwd, int64_t* __restrict__ pos,
                            unsigned long long* __restrict__ flags) {
  const int64_t j = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (j >= n3) return;
  const int64_t t = j / 3, s = j - 3 * t;
  const int64_t a = ids[6 * t + s], b = ids[6 * t + (s == 2 ? 0 : s + 1)];
  ka[j] = a < b ? a : b;
  kb[j] = a < b ? b : a;
  fwd[j] = a < b;
  pos[j] = j;
  wave_or(flags, a == b ? (unsigned)MHS_REFUSE_DEGENERATE : 0u);
}

// sort key of a column in the current order: its offset from the column's
// minimum (exact in unsigned arithmetic), so the sort needs only the span's bits
__global__ void k_gather_key(const int64_t* __restrict__ col, const int64_t* __restrict__ perm, int64_t n,
                             unsigned long long base, unsigned long long* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (i < n) out[i] = (unsigned long long)col[perm[i]] - base;
}

struct Cols {
  const int64_t* c[4];
  int n;
};

// flag[i] = 1 where the sorted key row i differs from row i - 1
__global__ void k_new_key(Cols cols, const int64_t* __restrict__ perm, int64_t n, int64_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (i >= n) return;
  int64_t f = i == 0;
  if (i > 0) {
    const int64_t p = perm[i], q = perm[i - 1];
    for (int c = 0; c < cols.n; ++c) f |= cols.c[c][p] != cols.c[c][q];
  }
  flag[i] = f;
}

// id of every occurrence, and the start of every run
__global__ void k_runs(const int64_t* __restrict__ flag, const int64_t* __restrict__ scan,
                       const int64_t* __restrict__ perm, int64_t n, int64_t* __restrict__ id_of,
                       int64_t id_base, int64_t id_stride, int64_t id_off, int64_t* __restrict__ start) {
  const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (i >= n) return;
  const int64_t u = scan[i] - 1;
  const int64_t j = perm[i];
  // occurrence j of the flat (triangle, slot) list -> ids[t * id_stride + id_off + slot]
  const int64_t t = j / 3, s = j - 3 * t;
  id_of[t * id_stride + id_off + s] = id_base + u;
  if (flag[i]) start[u] = i;
  if (i == n - 1) start[u + 1] = n;
}

// welded vertex = mean of its corners, summed in flat order
__global__ void k_vertex_mean(const double* __restrict__ tris, const int64_t* __restrict__ perm,
                              const int64_t* __restrict__ start, const int64_t* __restrict__ n_v_dev,
                              double* __restrict__ verts) {
  const int64_t v = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (v >= *n_v_dev) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  const int64_t a = start[v], b = start[v + 1];
  for (int64_t i = a; i < b; ++i) {
    const int64_t j = perm[i];
    sx = sx + tris[3 * j];
    sy = sy + tris[3 * j + 1];
    sz = sz + tris[3 * j + 2];
  }
  const double c = (double)(b - a);
  verts[3 * v] = sx / c;
  verts[3 * v + 1] = sy / c;
  verts[3 * v + 2] = sz / c;
}

// edge table: end points, multiplicity, forward traversals, midpoint node
__global__ void k_edge_rows(const int64_t* __restrict__ ka, const int64_t* __restrict__ kb,
                            const int* __restrict__ fwd, const int64_t* __restrict__ perm,
                            const int64_t* __restrict__ start, const int64_t* __restrict__ n_e_dev,
                            const double* __restrict__ verts, int64_t n_v, int64_t* __restrict__ edges,
                            int* __restrict__ counts, int* __restrict__ n_forward, double* __restrict__ coords,
                            unsigned long long* __restrict__ flags) {
  const int64_t e = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (e >= *n_e_dev) return;
  const int64_t a = start[e], b = start[e + 1];
  const int64_t j0 = perm[a];
  const int64_t p = ka[j0], q = kb[j0];
  int nf = 0;
  for (int64_t i = a; i < b; ++i) nf += fwd[perm[i]];
  const int cnt = (int)(b - a);
  edges[2 * e] = p;
  edges[2 * e + 1] = q;
  counts[e] = cnt;
  n_forward[e] = nf;
  for (int c = 0; c < 3; ++c) coords[3 * (n_v + e) + c] = 0.5 * (verts[3 * p + c] + verts[3 * q + c]);
  wave_or(flags, (cnt > 2 ? (unsigned)MHS_REFUSE_NON_MANIFOLD : 0u) | (cnt == 2 ? (unsigned)MHS_REFUSE_SHARED : 0u) |
                     (cnt == 2 && nf != 1 ? (unsigned)MHS_REFUSE_WINDING : 0u));
}

__global__ void k_vertex_coords(const double* __restrict__ verts, const int64_t* __restrict__ n_v_dev,
                                double* __restrict__ coords) {
  const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (i < 3 * *n_v_dev) coords[i] = verts[i];
}

// per node: area-weighted normal (normalised), edge weight, group of the
// last occurrence -- runs of the vertex sort, then of the edge sort
__global__ void k_node_sums(const int64_t* __restrict__ vperm, const int64_t* __restrict__ vstart,
                            const int64_t* __restrict__ eperm, const int64_t* __restrict__ estart, int64_t n_v,
                            int64_t n_e, const double* __restrict__ an, const double* __restrict__ area3,
                            const int64_t* __restrict__ groups, double* __restrict__ normals,
                            double* __restrict__ weights, int64_t* __restrict__ group) {
  const int64_t node = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (node >= n_v + n_e) return;
  const bool edge = node >= n_v;
  const int64_t r = edge ? node - n_v : node;
  const int64_t* perm = edge ? eperm : vperm;
  const int64_t a = (edge ? estart : vstart)[r], b = (edge ? estart : vstart)[r + 1];
  double nx = 0.0, ny = 0.0, nz = 0.0, w = 0.0;
  int64_t t = 0;
  for (int64_t i = a; i < b; ++i) {
    t = perm[i] / 3;
    nx = nx + an[3 * t];
    ny = ny + an[3 * t + 1];
    nz = nz + an[3 * t + 2];
    if (edge) w = w + area3[t];
  }
  const double norm = sqrt(nx * nx + ny * ny + nz * nz);
  normals[3 * node] = nx / norm;
  normals[3 * node + 1] = ny / norm;
  normals[3 * node + 2] = nz / norm;
  weights[node] = w;
  group[node] = groups[t];
}

__global__ void k_fill_u8(uint8_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (i < n) p[i] = 0;
}

// rim edges (one triangle) and their end vertices are the boundary
__global__ void k_boundary(const int64_t* __restrict__ edges, const int* __restrict__ counts, int64_t n_v,
                           int64_t n_e, uint8_t* __restrict__ boundary) {
  const int64_t e = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (e >= n_e || counts[e] != 1) return;
  boundary[n_v + e] = 1;
  boundary[edges[2 * e]] = 1;
  boundary[edges[2 * e + 1]] = 1;
}

// workspace layout, in 8-byte words after 256-byte alignment of each part
struct Work {
  int64_t *k[4], *tmp_key, *tmp_key2, *perm, *perm2, *flag, *scan, *vperm, *vstart, *eperm, *estart, *ka, *kb;
  int64_t *count_dev;         // [0] = n_v, [1] = n_e, [2] = refusal bits, [8..15] weld key ranges
  unsigned long long* partial;  // per-workgroup key ranges of k_weld_keys
  int* fwd;
  double* verts;
  void* sort_tmp;
  size_t sort_bytes;
  size_t total;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// rocPRIM's temporary storage for n keys: the larger of the sort's and the scan's
hipError_t prim_bytes(int64_t n, size_t& bytes) {
  size_t s = 0, c = 0;
  int64_t* none = nullptr;
  unsigned long long* ukey = nullptr;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, s, ukey, ukey, none, none, (unsigned)n, 0u, 64u);
  if (e != hipSuccess) return e;
  e = rocprim::inclusive_scan(nullptr, c, none, none, (size_t)n, rocprim::plus<int64_t>());
  bytes = s > c ? s : c;
  return e;
}

hipError_t layout(int64_t n_tri, uint8_t* base, Work& w) {
  const int64_t n3 = 3 * n_tri;
  const size_t v = align256((size_t)n3 * 8), v1 = align256((size_t)(n3 + 1) * 8);
  size_t off = 0;
  auto take = [&](size_t b) { uint8_t* p = base ? base + off : nullptr; off += b; return p; };
  for (int c = 0; c < 4; ++c) w.k[c] = (int64_t*)take(v);
  w.tmp_key = (int64_t*)take(v);
  w.tmp_key2 = (int64_t*)take(v);
  w.perm = (int64_t*)take(v);
  w.perm2 = (int64_t*)take(v);
  w.flag = (int64_t*)take(v);
  w.scan = (int64_t*)take(v);
  w.vperm = (int64_t*)take(v);
  w.vstart = (int64_t*)take(v1);
  w.eperm = (int64_t*)take(v);
  w.estart = (int64_t*)take(v1);
  w.ka = (int64_t*)take(v);
  w.kb = (int64_t*)take(v);
  w.count_dev = (int64_t*)take(256);
  w.partial = (unsigned long long*)take(align256((size_t)blocks_for(n3) * 64));
  w.fwd = (int*)take(align256((size_t)n3 * 4));
  w.verts = (double*)take(align256((size_t)n3 * 24));
  hipError_t e = prim_bytes(n3, w.sort_bytes);
  if (e != hipSuccess) return e;
  w.sort_tmp = take(align256(w.sort_bytes));
  w.total = off;
  return hipSuccess;
}

// Significant bits of a column's span (0: every key equal, no sort needed).
int span_bits(unsigned long long span) { return span ? 64 - __builtin_clzll(span) : 0; }

// Stable LSD sort of positions 0..n-1 by the key columns (most significant
// first in cols[]), each column sorted on its offset from base[c] over its
// bits[c] significant bits; leaves the sorted positions in out_perm and the
// runs' "new key" flags and inclusive scan in w.flag / w.scan.
int sort_unique(Work& w, int64_t** cols, const unsigned long long* base, const int* bits, int ncol, int64_t n,
                int64_t* out_perm, hipStream_t st) {
  int64_t *cur = w.perm, *nxt = w.perm2;     // w.perm holds 0..n-1 on entry
  unsigned long long* key_in = (unsigned long long*)w.tmp_key;
  unsigned long long* key_out = (unsigned long long*)w.tmp_key2;
  for (int c = ncol - 1; c >= 0; --c) {
    if (bits[c] == 0) continue;
    hipLaunchKernelGGL(k_gather_key, dim3(blocks_for(n)), dim3(TB), 0, st, cols[c], cur, n, base[c], key_in);
    size_t bytes = w.sort_bytes;
    if (mhs_internal_check(rocprim::radix_sort_pairs(w.sort_tmp, bytes, key_in, key_out, cur, nxt, (unsigned)n,
                                                     0u, (unsigned)bits[c], st),
                           "weld radix sort"))
      return -1;
    int64_t* s = cur; cur = nxt; nxt = s;
  }
  if (mhs_internal_check(hipMemcpyAsync(out_perm, cur, (size_t)n * 8, hipMemcpyDeviceToDevice, st),
                         "weld permutation copy"))
    return -1;
  Cols cc{};
  for (int c = 0; c < ncol; ++c) cc.c[c] = cols[c];
  cc.n = ncol;
  hipLaunchKernelGGL(k_new_key, dim3(blocks_for(n)), dim3(TB), 0, st, cc, out_perm, n, w.flag);
  size_t bytes = w.sort_bytes;
  return mhs_internal_check(rocprim::inclusive_scan(w.sort_tmp, bytes, w.flag, w.scan, (size_t)n,
                                                    rocprim::plus<int64_t>(), st),
                            "weld scan");
}

}  // namespace

extern "C" {

int64_t mhs_weld_workspace_bytes(int64_t n_tri) {
  if (n_tri <= 0 || 3 * n_tri > 0x7fffffffLL) return -1;
  Work w;
  if (layout(n_tri, nullptr, w) != hipSuccess) return -1;
  return (int64_t)w.total;
}

int mhs_shared_nodes(const double* tris, const int64_t* groups, int64_t n_tri, double tol,
                     const double* area_normals, const double* area3, void* work, int64_t work_bytes,
                     int64_t* ids, double* coords, double* normals, double* weights, int64_t* group,
                     uint8_t* boundary, int64_t* edges, int* counts, int* n_forward, int64_t* n_nodes,
                     int* refusals, void* stream) {
  if (!tris || !groups || !area_normals || !area3 || !work || !ids || !coords || !normals || !weights ||
      !group || !boundary || !edges || !counts || !n_forward || !n_nodes || !refusals)
    return mhs_internal_fail("mhs_shared_nodes: null pointer");
  if (n_tri <= 0 || 3 * n_tri > 0x7fffffffLL) return mhs_internal_fail("mhs_shared_nodes: n_tri out of range");
  if (!(tol > 0.0)) return mhs_internal_fail("mhs_shared_nodes: tol must be positive");
  hipStream_t st = (hipStream_t)stream;
  Work w;
  if (mhs_internal_check(layout(n_tri, (uint8_t*)work, w), "mhs_shared_nodes workspace")) return -1;
  if ((int64_t)w.total > work_bytes) return mhs_internal_fail("mhs_shared_nodes: workspace too small");
  const int64_t n3 = 3 * n_tri;
  const dim3 g3(blocks_for(n3)), blk(TB);

  // 1. weld: unique (kx, ky, kz, group) rows of the corners
  unsigned long long* range = (unsigned long long*)&w.count_dev[8];
  hipLaunchKernelGGL(k_weld_keys, g3, blk, 0, st, tris, groups, n3, tol, w.k[0], w.k[1], w.k[2], w.k[3], w.perm,
                     w.partial);
  hipLaunchKernelGGL(k_range_reduce, dim3(1), blk, 0, st, w.partial, (int64_t)g3.x, range);
  unsigned long long rh[8];
  if (mhs_internal_check(hipMemcpyAsync(rh, range, 64, hipMemcpyDeviceToHost, st), "weld key ranges") ||
      mhs_internal_check(hipStreamSynchronize(st), "mhs_shared_nodes sync"))
    return -1;
  unsigned long long wbase[4];
  int wbits[4];
  for (int c = 0; c < 4; ++c) {
    wbase[c] = rh[2 * c] ^ 0x8000000000000000ULL;      // the column's minimum, as raw int64 bits
    wbits[c] = span_bits(rh[2 * c + 1] - rh[2 * c]);
  }
  if (sort_unique(w, w.k, wbase, wbits, 4, n3, w.vperm, st)) return -1;
  hipLaunchKernelGGL(k_runs, g3, blk, 0, st, w.flag, w.scan, w.vperm, n3, ids, (int64_t)0, (int64_t)6,
                     (int64_t)0, w.vstart);
  if (mhs_internal_check(hipMemcpyAsync(&w.count_dev[0], &w.scan[n3 - 1], 8, hipMemcpyDeviceToDevice, st),
                         "weld count"))
    return -1;
  hipLaunchKernelGGL(k_vertex_mean, g3, blk, 0, st, tris, w.vperm, w.vstart, &w.count_dev[0], w.verts);
  hipLaunchKernelGGL(k_vertex_coords, dim3(blocks_for(3 * n3)), blk, 0, st, w.verts, &w.count_dev[0], coords);
  int64_t n_v = 0;
  if (mhs_internal_check(hipMemcpyAsync(&n_v, &w.count_dev[0], 8, hipMemcpyDeviceToHost, st), "vertex count") ||
      mhs_internal_check(hipStreamSynchronize(st), "mhs_shared_nodes sync"))
    return -1;

  // 2. edges: unique (min, max) vertex-id pairs of the three slots, ids in [0, n_v)
  unsigned long long* flags = (unsigned long long*)&w.count_dev[2];
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, flags, (int64_t)1);
  hipLaunchKernelGGL(k_edge_keys, g3, blk, 0, st, ids, n3, w.ka, w.kb, w.fwd, w.perm, flags);
  int64_t* ecols[2] = {w.ka, w.kb};
  const unsigned long long ebase[2] = {0, 0};
  const int ebits[2] = {span_bits((unsigned long long)(n_v - 1)), span_bits((unsigned long long)(n_v - 1))};
  if (sort_unique(w, ecols, ebase, ebits, 2, n3, w.eperm, st)) return -1;
  int64_t n_e = 0;
  if (mhs_internal_check(hipMemcpyAsync(&n_e, &w.scan[n3 - 1], 8, hipMemcpyDeviceToHost, st), "edge count") ||
      mhs_internal_check(hipStreamSynchronize(st), "mhs_shared_nodes sync"))
    return -1;
  if (mhs_internal_check(hipMemcpyAsync(&w.count_dev[1], &w.scan[n3 - 1], 8, hipMemcpyDeviceToDevice, st),
                         "edge count"))
    return -1;
  hipLaunchKernelGGL(k_runs, g3, blk, 0, st, w.flag, w.scan, w.eperm, n3, ids, n_v, (int64_t)6, (int64_t)3,
                     w.estart);
  hipLaunchKernelGGL(k_edge_rows, dim3(blocks_for(n_e)), blk, 0, st, w.ka, w.kb, w.fwd, w.eperm, w.estart,
                     &w.count_dev[1], w.verts, n_v, edges, counts, n_forward, coords, flags);

  // 3. per-node sums and the boundary
  hipLaunchKernelGGL(k_node_sums, dim3(blocks_for(n_v + n_e)), blk, 0, st, w.vperm, w.vstart, w.eperm, w.estart,
                     n_v, n_e, area_normals, area3, groups, normals, weights, group);
  hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(n_v + n_e)), blk, 0, st, boundary, n_v + n_e);
  hipLaunchKernelGGL(k_boundary, dim3(blocks_for(n_e)), blk, 0, st, edges, counts, n_v, n_e, boundary);
  if (mhs_internal_check(hipGetLastError(), "mhs_shared_nodes launch")) return -1;
  unsigned long long fl = 0;
  if (mhs_internal_check(hipMemcpyAsync(&fl, flags, 8, hipMemcpyDeviceToHost, st), "refusal flags") ||
      mhs_internal_check(hipStreamSynchronize(st), "mhs_shared_nodes sync"))
    return -1;
  n_nodes[0] = n_v;
  n_nodes[1] = n_e;
  refusals[0] = (int)fl;
  return 0;
}

int mhs_tri_geometry(const double* tris, int64_t n_tri, double* area_normals, double* area3, void* stats_dev,
                     double* stats, void* stream) {
  if (!tris || !area_normals || !area3 || !stats_dev || !stats) return mhs_internal_fail("mhs_tri_geometry: null pointer");
  if (n_tri <= 0 || n_tri > 0x7fffffffLL) return mhs_internal_fail("mhs_tri_geometry: n_tri out of range");
  hipStream_t st = (hipStream_t)stream;
  unsigned long long* sd = (unsigned long long*)stats_dev;
  hipLaunchKernelGGL(k_stats_init, dim3(1), dim3(64), 0, st, sd);
  hipLaunchKernelGGL(k_tri_geometry, dim3(blocks_for(n_tri)), dim3(TB), 0, st, tris, n_tri, area_normals, area3,
                     sd);
  if (mhs_internal_check(hipGetLastError(), "mhs_tri_geometry launch")) return -1;
  unsigned long long h[4];
  if (mhs_internal_check(hipMemcpyAsync(h, sd, 32, hipMemcpyDeviceToHost, st), "geometry stats") ||
      mhs_internal_check(hipStreamSynchronize(st), "mhs_tri_geometry sync"))
    return -1;
  double d0, d1;
  memcpy(&d0, &h[0], 8);
  memcpy(&d1, &h[1], 8);
  stats[0] = d0;
  stats[1] = d1;
  stats[2] = (double)h[2];
  stats[3] = (double)h[3];
  return 0;
}

}  // extern "C"
