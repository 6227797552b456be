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
// Memory: the caller passes every output at its capacity (3 n_tri vertices,
// 3 n_tri edges) and one workspace of mhs_weld_workspace_bytes(n_tri) bytes;
// nothing is allocated here.  The call synchronises its stream once, to read
// the two node counts.
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

__global__ void k_weld_keys(const double* __restrict__ tris, const int64_t* __restrict__ groups, int64_t n3,
                            double tol, int64_t* __restrict__ kx, int64_t* __restrict__ ky,
                            int64_t* __restrict__ kz, int64_t* __restrict__ kg, int64_t* __restrict__ pos) {
  const int64_t j = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (j >= n3) return;
  // np.round (half to even) of the IEEE quotient, then the exact conversion
  kx[j] = (int64_t)rint(tris[3 * j] / tol);
  ky[j] = (int64_t)rint(tris[3 * j + 1] / tol);
  kz[j] = (int64_t)rint(tris[3 * j + 2] / tol);
  kg[j] = groups[j / 3];
  pos[j] = j;
}

// local edge slots (v0 v1) (v1 v2) (v2 v0) per triangle, as (min, max); the
// welded vertex ids are columns 0..2 of the (n_tri, 6) id table
__global__ void k_edge_keys(const int64_t* __restrict__ ids, int64_t n3, int64_t* __restrict__ ka,
                            int64_t* __restrict__ kb, int* __restrict__ fwd, int64_t* __restrict__ pos) {
  const int64_t j = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (j >= n3) return;
  const int64_t t = j / 3, s = j - 3 * t;
  const int64_t a = ids[6 * t + s], b = ids[6 * t + (s == 2 ? 0 : s + 1)];
  ka[j] = a < b ? a : b;
  kb[j] = a < b ? b : a;
  fwd[j] = a < b;
  pos[j] = j;
}

__global__ void k_gather_i64(const int64_t* __restrict__ col, const int64_t* __restrict__ perm, int64_t n,
                             int64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (i < n) out[i] = col[perm[i]];
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
                            int* __restrict__ counts, int* __restrict__ n_forward, double* __restrict__ coords) {
  const int64_t e = (int64_t)blockIdx.x * TB + threadIdx.x;
  if (e >= *n_e_dev) return;
  const int64_t a = start[e], b = start[e + 1];
  const int64_t j0 = perm[a];
  const int64_t p = ka[j0], q = kb[j0];
  int nf = 0;
  for (int64_t i = a; i < b; ++i) nf += fwd[perm[i]];
  edges[2 * e] = p;
  edges[2 * e + 1] = q;
  counts[e] = (int)(b - a);
  n_forward[e] = nf;
  for (int c = 0; c < 3; ++c) coords[3 * (n_v + e) + c] = 0.5 * (verts[3 * p + c] + verts[3 * q + c]);
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
  int64_t *count_dev;         // [0] = n_v, [1] = n_e
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
  hipError_t e = rocprim::radix_sort_pairs(nullptr, s, none, none, none, none, (unsigned)n, 0u, 64u);
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
  w.fwd = (int*)take(align256((size_t)n3 * 4));
  w.verts = (double*)take(align256((size_t)n3 * 24));
  hipError_t e = prim_bytes(n3, w.sort_bytes);
  if (e != hipSuccess) return e;
  w.sort_tmp = take(align256(w.sort_bytes));
  w.total = off;
  return hipSuccess;
}

// Stable LSD sort of positions 0..n-1 by the key columns (most significant
// first in cols[]), leaving the sorted positions in out_perm and the runs'
// "new key" flags and inclusive scan in w.flag / w.scan.
int sort_unique(Work& w, int64_t** cols, int ncol, int64_t n, int64_t* out_perm, hipStream_t st) {
  int64_t *cur = w.perm, *nxt = w.perm2;     // w.perm holds 0..n-1 on entry
  for (int c = ncol - 1; c >= 0; --c) {
    hipLaunchKernelGGL(k_gather_i64, dim3(blocks_for(n)), dim3(TB), 0, st, cols[c], cur, n, w.tmp_key);
    size_t bytes = w.sort_bytes;
    if (mhs_internal_check(rocprim::radix_sort_pairs(w.sort_tmp, bytes, w.tmp_key, w.tmp_key2, cur, nxt,
                                                     (unsigned)n, 0u, 64u, st),
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
                     void* stream) {
  if (!tris || !groups || !area_normals || !area3 || !work || !ids || !coords || !normals || !weights ||
      !group || !boundary || !edges || !counts || !n_forward || !n_nodes)
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
  hipLaunchKernelGGL(k_weld_keys, g3, blk, 0, st, tris, groups, n3, tol, w.k[0], w.k[1], w.k[2], w.k[3], w.perm);
  if (sort_unique(w, w.k, 4, n3, w.vperm, st)) return -1;
  hipLaunchKernelGGL(k_runs, g3, blk, 0, st, w.flag, w.scan, w.vperm, n3, ids, (int64_t)0, (int64_t)6,
                     (int64_t)0, w.vstart);
  if (mhs_internal_check(hipMemcpyAsync(&w.count_dev[0], &w.scan[n3 - 1], 8, hipMemcpyDeviceToDevice, st),
                         "weld count"))
    return -1;
  hipLaunchKernelGGL(k_vertex_mean, g3, blk, 0, st, tris, w.vperm, w.vstart, &w.count_dev[0], w.verts);
  hipLaunchKernelGGL(k_vertex_coords, dim3(blocks_for(3 * n3)), blk, 0, st, w.verts, &w.count_dev[0], coords);

  // 2. edges: unique (min, max) vertex-id pairs of the three slots
  hipLaunchKernelGGL(k_edge_keys, g3, blk, 0, st, ids, n3, w.ka, w.kb, w.fwd, w.perm);
  int64_t* ecols[2] = {w.ka, w.kb};
  if (sort_unique(w, ecols, 2, n3, w.eperm, st)) return -1;
  if (mhs_internal_check(hipMemcpyAsync(&w.count_dev[1], &w.scan[n3 - 1], 8, hipMemcpyDeviceToDevice, st),
                         "edge count"))
    return -1;
  int64_t nn[2];
  if (mhs_internal_check(hipMemcpyAsync(nn, w.count_dev, 16, hipMemcpyDeviceToHost, st), "node counts") ||
      mhs_internal_check(hipStreamSynchronize(st), "mhs_shared_nodes sync"))
    return -1;
  const int64_t n_v = nn[0], n_e = nn[1];
  hipLaunchKernelGGL(k_runs, g3, blk, 0, st, w.flag, w.scan, w.eperm, n3, ids, n_v, (int64_t)6, (int64_t)3,
                     w.estart);
  hipLaunchKernelGGL(k_edge_rows, dim3(blocks_for(n_e)), blk, 0, st, w.ka, w.kb, w.fwd, w.eperm, w.estart,
                     &w.count_dev[1], w.verts, n_v, edges, counts, n_forward, coords);

  // 3. per-node sums and the boundary
  hipLaunchKernelGGL(k_node_sums, dim3(blocks_for(n_v + n_e)), blk, 0, st, w.vperm, w.vstart, w.eperm, w.estart,
                     n_v, n_e, area_normals, area3, groups, normals, weights, group);
  hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(n_v + n_e)), blk, 0, st, boundary, n_v + n_e);
  hipLaunchKernelGGL(k_boundary, dim3(blocks_for(n_e)), blk, 0, st, edges, counts, n_v, n_e, boundary);
  if (mhs_internal_check(hipGetLastError(), "mhs_shared_nodes launch")) return -1;
  n_nodes[0] = n_v;
  n_nodes[1] = n_e;
  return 0;
}

}  // extern "C"
