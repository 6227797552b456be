// gfx950 kernels of the mollified Mindlin half-space assembler and the
// C-ABI they are launched through (declared in include/mhs_amd.h).
//
// Launch plan of one matrix call (reference: src/mhs/kernels/assemble.py),
// observer rows in chunks, a 2-D grid (row, column or source tile) per kernel:
//   interaction with one or two components (contracted):
//     k_far   the far field (direct, image R, image Q) as contracted
//             pointwise series per (row, source) pair, all its node columns
//             from one evaluation of each point, written;
//     k_near  the closed-form near field and the folded mirror of the pairs
//             a row tile lists as having one, added;
//   every other form:
//     k_pair  direct + image R-family (or the folded mirror) + eigenstress,
//             written in the output's layout (displacement: row-tile lists,
//             far pairs first);
//     k_qm    the image Q-family, moments first, added;
//   then k_fold_list / k_fold_prep / k_fold (near and far points): the
//   Boussinesq surface integral of the folded pairs, added; k_gather the
//   continuous layout's shared-column sums; k_strain the elastic strain.
//   k_source_major copies a finished interaction matrix to the consumers'
//   source-major layout (its own entry point).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#define MHS_MAX_DEVICES 64   // devices a process may use the library on

#include "fold.hpp"
#include "../../include/mhs_amd.h"

#ifndef MHS_FAR_NPT32
#define MHS_FAR_NPT32 4   // quadrature points per lane of the far pass, FP32 points
#endif
#ifndef MHS_FAR_NPT64
#define MHS_FAR_NPT64 2   // quadrature points per lane of the far pass, FP64 points (4 spilled: 361 against 172 ms at n = 10^4)
#endif
// Lanes per pair of the far pass: fewer lanes put more pairs in a wave, so
// the per-pair set-up is issued once for more pairs, until the per-block
// kappa tables in LDS cut the occupancy (measured at n = 10^4: 8 lanes beat
// 16 by 13 % with FP32 points and by 20 % with FP64 points, both at 2 waves
// per SIMD; 4 lanes lose with FP32 points, LDS-limited to 2 waves per SIMD).
// With FP64 points 16 lanes halve the per-block kappa tables (79 -> 40 KB),
// so 3 waves per SIMD fit: FP64 far pass 183.7 -> 175.4 ms (4 waves: 188).
#ifndef MHS_FAR_LPP32
#define MHS_FAR_LPP32 8   // FP32 points, one output
#endif
#ifndef MHS_FAR_LPP64
#define MHS_FAR_LPP64 16  // FP64 points, one output
#endif
#ifndef MHS_FAR_LPP32B
#define MHS_FAR_LPP32B 16 // FP32 points, two outputs
#endif
#ifndef MHS_FAR_LPP64B
#define MHS_FAR_LPP64B 32 // FP64 points, two outputs
#endif
#ifndef MHS_FAR_WAVES32
#define MHS_FAR_WAVES32 3 // occupancy target of the far pass, FP32 points (waves per SIMD)
#endif
#ifndef MHS_FAR_WAVES64
#define MHS_FAR_WAVES64 3 // occupancy target of the far pass, FP64 points
#endif
// Occupancy targets of the closed-form kernels, whose per-item work is a
// called function: 2 waves per SIMD beat 1 (latency of the long FP64 chains)
// and 3 or 4 (spills), measured at n = 10^4: fold near points 47 -> 37 ms,
// near pass 18 -> 12.7 ms; measured again on the current kernels (mixed
// rule): 1 wave ran the near pass in 5.5 against 3.3 ms and the fold in
// 22.3 against 17.9 ms; 3 waves ran the fold in 22.6 against 21.1 ms (a
// build before the fold's own switch).
#ifndef MHS_FOLD_NEAR_WAVES
#define MHS_FOLD_NEAR_WAVES 2  // the fold's near-point kernel
#endif
// Threads per near-pass block: 256 (four waves), or 64 (one wave).  At
// n = 10^4 under the mixed rule's 2 L switch a row tile holds a few dozen
// near pairs, and one-wave blocks ran the near pass in 2.41 against 3.30 ms
// (FP64 rule 7.33 against 8.45 ms), bitwise the same; but the served stream
// of 512 requests lost with them (mixed p99 3.6 against 2.65 s, mean 0.26
// against 0.20 s, two runs each, same frozen stream), so four waves stay
// the default.  The queue form always takes four-wave blocks.
#ifndef MHS_NEAR_THREADS
#define MHS_NEAR_THREADS 256
#endif
#ifndef MHS_NEAR_WAVES
#define MHS_NEAR_WAVES 2       // the near pass
#endif
#ifndef MHS_PAIR_WAVES
#define MHS_PAIR_WAVES 2       // the pair pass of the displacement form
#endif
// The fold's far-point kernel (Gauss tables, or the mixed rule's direct
// series) at 3 waves per SIMD: at n = 10^4 its mixed-rule pass ran 1.7 ms
// faster than at 2 and 0.9 ms faster than at 4.
#ifndef MHS_FOLD_FAR_WAVES
#define MHS_FOLD_FAR_WAVES 3   // occupancy target of the fold's far-point kernel
#endif
#ifndef MHS_QC_CT
#define MHS_QC_CT 16   // default column tiles per block of the far pass
#endif

static thread_local char g_err[512];

static int fail(const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return -1;
}

static int check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    return -1;
  }
  return 0;
}

// The same message slot for the library's other translation unit (weld.hip).
__attribute__((visibility("hidden"))) int mhs_internal_fail(const char* msg) { return fail(msg); }
__attribute__((visibility("hidden"))) int mhs_internal_check(hipError_t e, const char* what) {
  return check(e, what);
}

// Blocks of 256 for a grid-stride loop over n items, capped so the dispatch's
// work-item count stays far inside 32 bits.
static unsigned grid_1d(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > (1 << 20) ? (1 << 20) : b));
}

// ------------------------------------------------------------ block order --
// The tile (bx, by) = (column tile, observer row) of a 2-D launch with the
// rows walked first: dispatch order runs x fastest, so the consecutive
// blocks of linear index x + y * gridDim.x take every row of one column tile
// before the next tile.  The blocks resident at once then read the same
// few thousand source records, which stay in each XCD's L2; with x as the
// tile they each read a different slice of all the records (at n = 10^4
// sources 6.4 MB, more than one XCD's L2).  Measured on the interaction at
// n = 10^4: mixed far pass 17.4 -> 14.9 ms, near pass 5.5 -> 3.6 ms; the
// three-fault network's mixed assembly 762 -> 713 ms.
#ifndef MHS_ROWS_FIRST
#define MHS_ROWS_FIRST 1
#endif
__device__ __forceinline__ void rows_first(int64_t& bx, int64_t& by) {
  const int64_t lin = blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
  if (MHS_ROWS_FIRST) {
    bx = lin / gridDim.y;
    by = lin - bx * gridDim.y;
  } else {
    bx = blockIdx.x;
    by = blockIdx.y;
  }
}

// ------------------------------------------------------------ work queue --
// A batch of assembly launches (several requests, meshes of uneven size)
// runs each tiled pass as ONE persistent launch per kernel instance: the
// tiles of every launch that needs that instance form a ragged queue of
// segments, segment g holding the nbx x rows tiles of one launch unit
// (parameter block p) from global tile index start on, in admission order.
// Each workgroup takes the next tile index with one atomic, finds its
// segment by binary search and runs the ordinary per-tile body on it, until
// the queue is empty: no workgroup ever waits on another, and every
// workgroup exits once the counter passes the total.
static_assert(sizeof(Params) % 4 == 0, "parameter blocks are copied as 32-bit words");

struct QueueSeg {
  int64_t start;   // first global tile index of the segment
  int32_t p;       // parameter block
  int32_t nbx;     // tiles per row (the grid's x extent)
};
struct QueueArgs {
  const Params* __restrict__ params;
  const QueueSeg* __restrict__ seg;
  unsigned long long* counter;       // zero before the launch
  int64_t total;                     // tiles over all segments
  int32_t n_seg;
};

// Segment of global tile b: (parameter block, bx, by).
__device__ __forceinline__ void queue_tile(const QueueArgs& q, long long b, int& p, int& bx, int& by) {
  int lo = 0, hi = q.n_seg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (q.seg[mid].start <= b) lo = mid;
    else hi = mid - 1;
  }
  const long long r = b - q.seg[lo].start;
  p = q.seg[lo].p;
  bx = (int)(r % q.seg[lo].nbx);
  by = (int)(r / q.seg[lo].nbx);
}

// Parameter block p, referenced in the constant address space as a kernel
// argument is: its fields are scalar loads (the address is uniform and the
// memory invariant during the launch), so the passes see them as uniform
// values, and non-inlined calls receive a pointer to global memory, never to
// a private copy.  (A private copy passed to such a call lived in scratch and
// its fields came back as vector loads: uniform loop exits became divergent,
// and the queue loop's compiled form then reloaded a register across its
// back edge from a spill slot that a block run with no active lane had not
// written -- a wild atomic address on the next tile.)
__device__ __forceinline__ const Params& queue_params(const QueueArgs& q, int p) {
  typedef const __attribute__((address_space(4))) Params* ConstParams;
  const ConstParams c = (ConstParams)(q.params + p);
  return *(const Params*)c;
}

template <class Body>
__device__ __forceinline__ void drain_queue(const QueueArgs& q, Body body) {
  __shared__ long long tile_s;
  __shared__ int p_s, bx_s, by_s;
  for (;;) {
    __syncthreads();                 // the previous tile's LDS is no longer read
    if (threadIdx.x == 0) {
      const long long b = (long long)atomicAdd(q.counter, 1ULL);
      tile_s = b;
      if (b < q.total) queue_tile(q, b, p_s, bx_s, by_s);
    }
    __syncthreads();
    if (tile_s >= q.total) return;   // the same in every lane
    const int p = __builtin_amdgcn_readfirstlane(p_s);
    const int64_t bx = __builtin_amdgcn_readfirstlane(bx_s), by = __builtin_amdgcn_readfirstlane(by_s);
    body(queue_params(q, p), bx, by);
  }
}



// ------------------------------------------------------------------ pair --
// Row tiles of the pair kernels: a block (256 threads) takes one receiver
// row and MHS_ROW_TILE items of it (columns, or sources) from c0 on, of n in
// all, and sorts them into two LDS lists by cls(item) (1: the first list,
// from list[0]; 2: the second, from list[MHS_ROW_TILE]; 0: neither), each in
// ascending order, by wave ballots; the lists then run with every lane busy
// and without mixing the two kinds of work in a wave.
#define MHS_ROW_TILE 2048

// skip(first column of a step) may declare a whole step of 256 columns
// class 0 without classifying them (the same answer in every thread of the
// block: it may depend on the block's row and the step only).
struct NoSkip {
  __device__ __forceinline__ bool operator()(int64_t) const { return false; }
};

template <typename Cls, typename Skip = NoSkip>
__device__ __forceinline__ void list_row_tile(int64_t n, int64_t c0, Cls cls, int* list, int& n1, int& n2,
                                              Skip skip = Skip()) {
  __shared__ int cnt_s[2][4];
  const int tid = threadIdx.x, wave = tid >> 6;
  n1 = n2 = 0;
  for (int step = 0; step < MHS_ROW_TILE; step += 256) {
    // a skipped step lists nothing: every thread of the block takes this
    // branch together (skip depends on the step alone), so no barrier runs
    if (skip(c0 + step)) {
      if (c0 + step + 256 >= n) break;
      continue;
    }
    const int64_t col = c0 + step + tid;
    const int c = col < n ? cls(col) : 0;
    const unsigned long long m1 = __ballot(c == 1), m2 = __ballot(c == 2);
    if ((tid & 63) == 0) {
      cnt_s[0][wave] = __popcll(m1);
      cnt_s[1][wave] = __popcll(m2);
    }
    __syncthreads();
    int o1 = n1, o2 = n2, s1 = 0, s2 = 0;
#pragma unroll
    for (int wv = 0; wv < 4; ++wv) {
      o1 += wv < wave ? cnt_s[0][wv] : 0;
      o2 += wv < wave ? cnt_s[1][wv] : 0;
      s1 += cnt_s[0][wv];
      s2 += cnt_s[1][wv];
    }
    if (c == 1)
      list[o1 + __builtin_amdgcn_mbcnt_hi((unsigned)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m1, 0u))] =
          step + tid;
    if (c == 2)
      list[MHS_ROW_TILE + o2 +
           __builtin_amdgcn_mbcnt_hi((unsigned)(m2 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m2, 0u))] =
          step + tid;
    n1 += s1;
    n2 += s2;
    __syncthreads();          // cnt_s is rewritten by the next step
    if (c0 + step + 256 >= n) break;
  }
}

// list_row_tile for a block of one wave (64 threads): the same two lists in
// the same order, steps of 64 columns, no barrier inside the loop (the
// caller's block barrier follows before the lists are read).
template <typename Cls, typename Skip = NoSkip>
__device__ __forceinline__ void list_row_tile_wave(int64_t n, int64_t c0, Cls cls, int* list, int& n1, int& n2,
                                                   Skip skip = Skip()) {
  const int tid = threadIdx.x;
  n1 = n2 = 0;
  for (int step = 0; step < MHS_ROW_TILE; step += 64) {
    if (!skip(c0 + step)) {
      const int64_t col = c0 + step + tid;
      const int c = col < n ? cls(col) : 0;
      const unsigned long long m1 = __ballot(c == 1), m2 = __ballot(c == 2);
      if (c == 1)
        list[n1 + __builtin_amdgcn_mbcnt_hi((unsigned)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m1, 0u))] =
            step + tid;
      if (c == 2)
        list[MHS_ROW_TILE + n2 +
             __builtin_amdgcn_mbcnt_hi((unsigned)(m2 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m2, 0u))] =
            step + tid;
      n1 += __popcll(m1);
      n2 += __popcll(m2);
    }
    if (c0 + step + 64 >= n) break;
  }
}

// One pair of the uncontracted forms as a function of its own (inlined into
// the list loop its closed forms spill more).
template <int P, int MODE>
__device__ __attribute__((noinline)) void pair_item_call(const Params& pr, int64_t o, int64_t col) {
  pair_item<P, MODE>(pr, o, col);
}

// The pair pass of the uncontracted forms.  Displacement: a row tile's pairs
// that are far on every part (Gauss tables only) first, then those with a
// closed-form part (direct near, image near, or the folded mirror), at two
// waves per SIMD.  The stress family keeps a thread per pair at one wave per
// SIMD: measured at n = 2000 (stress, traction, 3 x 3 interaction) the lists
// and the higher occupancy were 5-10 % slower for it, and 10-25 % faster
// for displacement.
template <int P, int MODE>   // MODE 0: stress family (eigenstress included), 1: displacement
__device__ __forceinline__ void body_pair(const Params& pr, int64_t bx_, int64_t by_) {
  constexpr int K = (P + 1) * (P + 2) / 2;
  if (MODE != 1) {
    const int64_t col = (int64_t)bx_ * 256 + threadIdx.x;
    if (col < pr.n_cols) pair_item<P, MODE>(pr, pr.obs_lo + by_, col);
    return;
  }
  __shared__ int list_s[2 * MHS_ROW_TILE];
  const int64_t o = pr.obs_lo + by_;
  const int64_t c0 = (int64_t)bx_ * MHS_ROW_TILE;
  const double ob[3] = {pr.obs[3 * o], pr.obs[3 * o + 1], pr.obs[3 * o + 2]};
  int n_far, n_near;
  list_row_tile(pr.n_cols, c0, [&](int64_t col) {
    return near_parts(pr, pr.src + (col / K) * MHS_SRC_STRIDE, ob) ? 2 : 1;
  }, list_s, n_far, n_near);
  for (int i = threadIdx.x; i < n_far + n_near; i += 256) {
    __asm__ volatile("" ::: "memory");   // no loads hoisted out of the pair loop
    const int c = i < n_far ? list_s[i] : list_s[MHS_ROW_TILE + i - n_far];
    pair_item_call<P, MODE>(pr, o, c0 + c);
  }
}
template <int P, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == 1 ? MHS_PAIR_WAVES : 1)))
k_pair(Params pr) {
  int64_t bx, by;
  rows_first(bx, by);
  body_pair<P, MODE>(pr, bx, by);
}
// the same pass over the tiles of a batch's queue (see drain_queue)
template <int P, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == 1 ? MHS_PAIR_WAVES : 1)))
kq_pair(QueueArgs q) {
  drain_queue(q, [&](const Params& pr, int64_t bx, int64_t by) { body_pair<P, MODE>(pr, bx, by); });
}

// --------------------------------------------------------------- Q-family --
// Q-family of the uncontracted forms (disp, stress, traction, interaction
// with more than two components), moments first: a thread per (observer row,
// element column); per chunk of monomials the budget-law Gauss points are
// summed into the chunk's moments w d1^a d2^b d3^c R^-n Q^-q, then the
// chunk's records are folded into the traction-operator tensor S
// (image.py:107-162); the stress (or displacement) from S is added.
template <int P, bool STRESS>
__device__ __forceinline__ void body_qm(const Params& pr, int64_t bx_, int64_t by_) {
  constexpr int K = (P + 1) * (P + 2) / 2;
  constexpr int NS = STRESS ? 27 : 9;
  constexpr int NCH = STRESS ? MHS_QMOM_DDG_CHUNKS : MHS_QMOM_DG_CHUNKS;
  constexpr int CH = STRESS ? MHS_QMOM_DDG_CHUNK : MHS_QMOM_DG_CHUNK;
  const int64_t o = pr.obs_lo + by_;
  const int64_t col = (int64_t)bx_ * 256 + threadIdx.x;
  if (col >= pr.n_cols) return;
  const int64_t s = col / K;
  const int k = (int)(col - s * K);
  const double* src = pr.src + s * MHS_SRC_STRIDE;
  const double ob[3] = {pr.obs[3 * o], pr.obs[3 * o + 1], pr.obs[3 * o + 2]};
  if (in_fold_band(src, ob, pr.flags)) return;        // the fold replaces R + Q there
  int nq;
  if (pr.mixed_rule) {
    // the mixed rule's order for a far pair, the budget law for a near one
    const double Di = centroid_distance(src + S_IC, ob, src[S_EPS]);
    nq = Di <= pr.near_r * src[S_IL] ? q_order_fast<P>(src, ob) : rule_order(pr.rule, Di, src[S_IL]) + P;
  } else {
    nq = q_order_fast<P>(src, ob);
  }
  const double* n = src + S_N;
  const double* v = src + S_V;
  const double eps2 = src[S_EPS] * src[S_EPS];
  const double ar2 = 2.0 * src[S_AREA];
  const int off = nq * (nq - 1) / 2;
  const double* gx = pr.gl_x + off;
  const double* gw = pr.gl_w + off;
  double S[NS];
#pragma unroll
  for (int t = 0; t < NS; ++t) S[t] = 0.0;
  // one chunk of monomials: moments over the points, then the records folded
  auto chunk = [&](int c) {
    double M[CH];
#pragma unroll
    for (int t = 0; t < CH; ++t) M[t] = 0.0;
    for (int i = 0; i < nq; ++i) {
      const double gi = gx[i], wi = gw[i];
      for (int j = 0; j < nq; ++j) {
        const double x1 = gi, x2 = gx[j] * (1.0 - gi);
        const double w = wi * gw[j] * (1.0 - gi) * ar2;
        const double l0 = 1.0 - x1 - x2;
        const double y0 = l0 * v[0] + x1 * v[3] + x2 * v[6];
        const double y1 = l0 * v[1] + x1 * v[4] + x2 * v[7];
        const double y2 = l0 * v[2] + x1 * v[5] + x2 * v[8];
        const double wn = w * shape_at<P>(k, l0, x1, x2);
        const double d1 = ob[0] - y0, d2 = ob[1] - y1, d3 = ob[2] + y2;
        const double R2 = d1 * d1 + d2 * d2 + d3 * d3 + eps2;
        const double ir = rsq_refined(R2), iq = rcp_refined(R2 * ir - d3);
        if (STRESS) qmom_ddg_acc(c, d1, d2, d3, ir, iq, wn, M);
        else qmom_dg_acc(c, d1, d2, d3, ir, iq, wn, M);
      }
    }
    if (STRESS) qmom_ddg_fold(c, M, pr.coef, n, pr.mu, pr.lam, S);
    else qmom_dg_fold(c, M, pr.coef, n, pr.mu, pr.lam, S);
  };
  // order 2, stress family: the chunk loop stays a loop (unrolled, the
  // chunks' folds are scheduled together and the point loop spills: 46 ->
  // 56 ms per served dispatch); every other instance unrolls it
  if (P == 2 && STRESS) {
#pragma unroll 1
    for (int c = 0; c < NCH; ++c) chunk(c);
  } else {
#pragma unroll
    for (int c = 0; c < NCH; ++c) chunk(c);
  }
  if (STRESS) {
    double H[27];
    stress_from_S(S, pr.mu, pr.lam, H);
    emit_stress(pr, o, col, s, H, 0.0, 0.0, n, n, true);
  } else {
    emit_disp(pr, o, col, S, true);
  }
}
template <int P, bool STRESS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_qm(Params pr) {
  int64_t bx, by;
  rows_first(bx, by);
  body_qm<P, STRESS>(pr, bx, by);
}
// (no queue form: compiled with its parameter block in memory instead of in
// kernel arguments, its point loop spills about 10 KB per lane; a batch
// launches it per range)


// ------------------------------------------- contracted interaction, far --
// Interaction with one or two output components (receiver frame row a,
// source frame row b).  Every far-field term of a pair is a quadrature sum of
// a pointwise kernel, so each is contracted with the pair's output functional
// once per pair and the point loop sums one scalar series:
//
//   direct (Kelvin-Cortez, source triangle, far when sqrt(|x-c|^2+eps^2) > 10 L):
//     -C1 sum w N [A3/R^3 + (A5c c + A5a a + A5b b + A5e)/R^5
//                  + (15 a b + A7c c + A7e)/R^7],
//     a = d.E.d, b = d.F.d, c = d.sym(EF).d, E the receiver half (traction
//     along frame row a on the receiver plane, in strain space), F the source
//     half (Hooke of slip along frame row b on the source plane), A7e the
//     subtracted eigenstress (15 eps^4 / 8 pi) rb.F.nr / C1;
//   image R-family (mirrored triangle, far by the same rule, not folded):
//     sum w N sum_k kappaR_k d^abc R^-n, kappaR_k = sum_v F_v BR[k][v];
//   image Q-family (not folded, budget-law order):
//     sum w N sum_k kappaQ_k d^abc R^-n Q^-q, kappaQ_k = sum_v F_v BQ[k][v];
//
// BR, BQ (the image records contracted with E, BR with the receiver-depth
// powers of its coefficients) are built once per block: a block is one
// receiver row and pr.far_tiles column tiles of PPB = 256 / LPP pairs.  LPP
// lanes per pair, NPT points per lane per chunk of LPP * NPT points; the
// per-pair set-up (classification, kappas, direct coefficients) runs once per
// 64 / LPP pairs of a wave.  The kappas live in LDS and every read feeds NPT
// FMAs.  T is the point arithmetic (double, or float in
// the mixed mode); point geometry, weights and all sums stay in double.  The
// near parts (closed forms, folded mirror) are added by k_near.
#define MHS_DC 25   // per-output direct coefficients: E(6) F(6) sym(EF)(6) A3 A5a A5b A5e A7c A7e A5c

template <typename T, int NPT, int NOUT>
__device__ __forceinline__ void d_blocked(const T* __restrict__ kd, const T* x, const T* y, const T* z,
                                          const T* ir, T* s) {
  T p00[NPT], p11[NPT], p22[NPT], p01[NPT], p02[NPT], p12[NPT], i2[NPT], i3[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    p00[j] = x[j] * x[j]; p11[j] = y[j] * y[j]; p22[j] = z[j] * z[j];
    p01[j] = (x[j] + x[j]) * y[j]; p02[j] = (x[j] + x[j]) * z[j]; p12[j] = (y[j] + y[j]) * z[j];
    i2[j] = ir[j] * ir[j];
    i3[j] = i2[j] * ir[j];
  }
#pragma unroll
  for (int o = 0; o < NOUT; ++o) {
    const T* c = kd + o * MHS_DC;
    T cf[MHS_DC];
#pragma unroll
    for (int t = 0; t < MHS_DC; ++t) cf[t] = c[t];
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      const T a = fma(cf[0], p00[j], fma(cf[1], p11[j], fma(cf[2], p22[j],
                  fma(cf[3], p01[j], fma(cf[4], p02[j], cf[5] * p12[j])))));
      const T b = fma(cf[6], p00[j], fma(cf[7], p11[j], fma(cf[8], p22[j],
                  fma(cf[9], p01[j], fma(cf[10], p02[j], cf[11] * p12[j])))));
      const T cc = fma(cf[12], p00[j], fma(cf[13], p11[j], fma(cf[14], p22[j],
                   fma(cf[15], p01[j], fma(cf[16], p02[j], cf[17] * p12[j])))));
      const T s5 = fma(cf[24], cc, fma(cf[19], a, fma(cf[20], b, cf[21])));
      const T s7 = fma((T)15 * a, b, fma(cf[22], cc, cf[23]));
      s[o * NPT + j] = fma(fma(s7, i2[j], s5), i2[j], cf[18]) * i3[j];
    }
  }
}

// 1/sqrt and 1/x of the point series: the hardware approximations (1 ulp)
// in FP32; in FP64 the hardware approximations refined by two Newton steps
// each (the seed's ~2^-22 relative error squares per step, to the last bits
// of a double), about half the instructions of a correctly rounded sqrt and
// two divisions.  R2 >= eps^2 > 0 and R - d3 >= R > 0 (d3 <= 0 on the
// mirrored side), so no special values arise.
__device__ __forceinline__ void radii(float R2, float d3, bool q, float& ir, float& iq) {
  ir = __builtin_amdgcn_rsqf(R2);
  iq = q ? __builtin_amdgcn_rcpf(R2 * ir - d3) : 0.0f;
}
__device__ __forceinline__ void radii(double R2, double d3, bool q, double& ir, double& iq) {
  ir = rsq_refined(R2);
  iq = q ? rcp_refined(R2 * ir - d3) : 0.0;
}

// One quadrature pass of a pair over the ng x ng collapsed Gauss points of
// the source triangle (IMG false: d = x - y) or of its mirror (IMG true:
// d = x - y*), summing the selected series, times wscale, into acc[k][t] for
// every node k of the source: the series of a point does not depend on the
// node, only its weight N_k does, so a point is evaluated once for all K
// nodes.  LPP lanes share the points, NPT per lane per chunk of LPP * NPT.
// The point comes from the flattened rule table; the offset is
// d = (x - v0) - x1 (v1 - v0) - x2 (v2 - v0) in FP64, mirrored in z for IMG.
template <typename T, int P, int NOUT, int NPT, int LPP, bool IMG, bool DIR, bool RF, bool QF>
__device__ __forceinline__ void far_pass(const Params& pr, const double* __restrict__ src,
                                         const double ob[3], int ng, int li, const T* kd, const T* kr,
                                         const T* kq, double wscale, double (*acc)[NOUT]) {
  constexpr int K = (P + 1) * (P + 2) / 2;
  constexpr int CH = LPP * NPT;
  const double* v = src + S_V;
  const double sz = IMG ? -1.0 : 1.0;           // z of the (mirrored) vertices
  const double ov[3] = {ob[0] - v[0], ob[1] - v[1], ob[2] - sz * v[2]};
  const double e1[3] = {v[3] - v[0], v[4] - v[1], sz * (v[5] - v[2])};
  const double e2[3] = {v[6] - v[0], v[7] - v[1], sz * (v[8] - v[2])};
  const T eps2 = (T)(src[S_EPS] * src[S_EPS]);
  const double ar2 = 2.0 * src[S_AREA] * wscale;
  const size_t toff = (size_t)(ng - 1) * ng * (2 * ng - 1) / 6;
  const double* qx1 = pr.gq + toff;
  const double* qx2 = qx1 + MHS_GQ_TOTAL;
  const double* qw = qx1 + 2 * MHS_GQ_TOTAL;
  const int npts = ng * ng;
  for (int base = 0; base < npts; base += CH) {
    // compiler-only fence: keeps the kappa reads inside the chunk instead of
    // hoisting all of them out of the loop into registers
    __asm__ volatile("" ::: "memory");
    T x[NPT], y[NPT], z[NPT], ir[NPT], iq[NPT], sv[NOUT * NPT];
    double wn[NPT];
    T b1[NPT], b2[NPT];                         // barycentric point (P > 0: node weights)
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
      const int g0 = base + u * LPP + li;
      const bool in = g0 < npts;
      const int g = in ? g0 : 0;
      const double x1 = qx1[g], x2 = qx2[g];
      wn[u] = in ? qw[g] * ar2 : 0.0;
      b1[u] = (T)x1;
      b2[u] = (T)x2;
      const T d1 = (T)fma(-x1, e1[0], fma(-x2, e2[0], ov[0]));
      const T d2 = (T)fma(-x1, e1[1], fma(-x2, e2[1], ov[1]));
      const T d3 = (T)fma(-x1, e1[2], fma(-x2, e2[2], ov[2]));
      x[u] = d1; y[u] = d2; z[u] = d3;
      radii(d1 * d1 + d2 * d2 + d3 * d3 + eps2, d3, QF, ir[u], iq[u]);
    }
    if (DIR) d_blocked<T, NPT, NOUT>(kd, x, y, z, ir, sv);
    if (RF) {
      __asm__ volatile("" ::: "memory");
      r_blocked<T, NPT, NOUT>(kr, x, y, z, ir, iq, sv);
    }
    if (QF) {
      __asm__ volatile("" ::: "memory");
      q_blocked<T, NPT, NOUT>(kq, x, y, z, ir, iq, sv);
    }
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
#pragma unroll
      for (int t = 0; t < NOUT; ++t) {
        if (P == 0) {
          acc[0][t] = fma(wn[u], (double)sv[t * NPT + u], acc[0][t]);
        } else {
          const double val = wn[u] * (double)sv[t * NPT + u];
          const T l1 = b1[u], l2 = b2[u], l0 = (T)1 - l1 - l2;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            T nk;
            if (P == 1) nk = k == 0 ? l0 : k == 1 ? l1 : l2;
            else if (k < 3) { const T l = k == 0 ? l0 : k == 1 ? l1 : l2; nk = l * ((T)2 * l - (T)1); }
            else nk = (T)4 * (k == 3 ? l0 * l1 : k == 4 ? l1 * l2 : l2 * l0);
            acc[k][t] = fma((double)nk, val, acc[k][t]);
          }
        }
      }
    }
  }
}

template <typename T, int P, int NOUT, int NPT, int LPP>
__device__ __forceinline__ void body_far(const Params& pr, int64_t bx_, int64_t by_) {
  static_assert(LPP >= 4 && LPP <= 64 && (LPP & (LPP - 1)) == 0, "lanes per pair: a power of two in [4, 64]");
  constexpr int K = (P + 1) * (P + 2) / 2;
  constexpr int PPB = 256 / LPP;
  constexpr int QB = MHS_QM * 6, RB = MHS_RM * 6;
  __shared__ double BQ_s[NOUT][QB];
  __shared__ double BR_s[NOUT][RB];
  __shared__ double E_s[NOUT][9];
  __shared__ T kq_s[PPB][NOUT * MHS_QM];
  __shared__ T kr_s[PPB][NOUT * MHS_RM];
  __shared__ T kd_s[PPB][NOUT * MHS_DC];
  __shared__ double res_s[PPB * K * NOUT];  // a tile's results, stored as one contiguous run
  const int tid = threadIdx.x;
  const int n_a = pr.n_a, n_b = pr.n_b;
  const int64_t o = pr.obs_lo + by_;
  const double ob[3] = {pr.obs[3 * o], pr.obs[3 * o + 1], pr.obs[3 * o + 2]};
  const double* nr = pr.nrm + 3 * o;
  // receiver half E_a (strain-space traction functional, frame row a)
  if (tid < n_a * 9) {
    const int a = tid / 9, i = (tid % 9) / 3, p = tid % 3;
    const double* rb = pr.rb + ((int64_t)o * pr.na_tot + a) * 3;
    const double rn = rb[0] * nr[0] + rb[1] * nr[1] + rb[2] * nr[2];
    E_s[a][i * 3 + p] = pr.mu * (rb[i] * nr[p] + rb[p] * nr[i]) + (i == p ? pr.lam * rn : 0.0);
  }
  __syncthreads();
  for (int t = tid; t < n_a * QB; t += 256) {
    const int a = t / QB, kv = t % QB;
    double acc = 0.0;
    for (int r = mhs_qb_ptr[kv]; r < mhs_qb_ptr[kv + 1]; ++r)
      acc = fma(pr.coef[mhs_qb_coef[r]], E_s[a][mhs_qb_ip[r]], acc);
    BQ_s[a][kv] = acc;
  }
  {
    const double zp[3] = {1.0, ob[2], ob[2] * ob[2]};
    for (int t = tid; t < n_a * RB; t += 256) {
      const int a = t / RB, kv = t % RB;
      double acc = 0.0;
      for (int r = mhs_rb_ptr[kv]; r < mhs_rb_ptr[kv + 1]; ++r)
        acc = fma(pr.coef[mhs_rb_coef[r]] * zp[mhs_rb_pz[r]], E_s[a][mhs_rb_ip[r]], acc);
      BR_s[a][kv] = acc;
    }
  }
  __syncthreads();
  for (int tile = 0; tile < pr.far_tiles; ++tile) {
    // the lane's pair and LDS addresses are derived afresh each tile (the
    // empty asm redefines the lane index), so they are not held, or spilled,
    // across the tile loop
    unsigned lane = tid;
    __asm__ volatile("" : "+v"(lane));
    lane &= 255u;
    const int grp = (int)(lane / LPP), li = (int)(lane % LPP);
    // a pair is (receiver row, source): its K node columns share every point
    const int64_t s = ((int64_t)bx_ * pr.far_tiles + tile) * PPB + grp;
    const bool valid = s < pr.n_src;
    int nd = 0, ni = 0, nq = 0;
    bool dfar = false, rfar = false, img = false;
    const double* src = pr.src;
    if (valid) {
      src = pr.src + s * MHS_SRC_STRIDE;
      const double eps = src[S_EPS];
      const double Dd = centroid_distance(src + S_C, ob, eps);
      dfar = !(Dd <= 10.0 * src[S_L]);
      nd = Dd > 40.0 * src[S_L] ? 8 : 12;
      img = !in_fold_band(src, ob, pr.flags);
      if (img) {
        const double Di = centroid_distance(src + S_IC, ob, eps);
        rfar = !(Di <= 10.0 * src[S_IL]);
        ni = Di > 40.0 * src[S_IL] ? 8 : 12;
        nq = q_order_fast<P>(src, ob);
      }
      // source half F_b (Hooke on the source plane, slip along frame row b)
      const double* n = src + S_N;
      double F[2][6];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (b < n_b) {
          const double* sb = pr.sb + (s * pr.nb_tot + b) * 3;
          const double sn = sb[0] * n[0] + sb[1] * n[1] + sb[2] * n[2];
          F[b][0] = 2.0 * pr.mu * sb[0] * n[0] + pr.lam * sn;
          F[b][1] = 2.0 * pr.mu * sb[1] * n[1] + pr.lam * sn;
          F[b][2] = 2.0 * pr.mu * sb[2] * n[2] + pr.lam * sn;
          F[b][3] = pr.mu * (sb[0] * n[1] + n[0] * sb[1]);
          F[b][4] = pr.mu * (sb[0] * n[2] + n[0] * sb[2]);
          F[b][5] = pr.mu * (sb[1] * n[2] + n[1] * sb[2]);
        } else {
#pragma unroll
          for (int v = 0; v < 6; ++v) F[b][v] = 0.0;
        }
      }
      if (img) {
        for (int t = li; t < NOUT * MHS_QM; t += LPP) {
          const int out = t / MHS_QM, km = t - out * MHS_QM;
          const int a = out / n_b, b = out - a * n_b;
          const double* Bk = &BQ_s[a][km * 6];
          const double* Fb = F[b == 0 ? 0 : 1];
          double kv = Fb[0] * Bk[0];
#pragma unroll
          for (int v = 1; v < 6; ++v) kv = fma(Fb[v], Bk[v], kv);
          kq_s[grp][t] = (T)kv;
        }
      }
      if (img && rfar) {
        for (int t = li; t < NOUT * MHS_RM; t += LPP) {
          const int out = t / MHS_RM, km = t - out * MHS_RM;
          const int a = out / n_b, b = out - a * n_b;
          const double* Bk = &BR_s[a][km * 6];
          const double* Fb = F[b == 0 ? 0 : 1];
          double kv = Fb[0] * Bk[0];
#pragma unroll
          for (int v = 1; v < 6; ++v) kv = fma(Fb[v], Bk[v], kv);
          kr_s[grp][t] = (T)kv;
        }
      }
      if (dfar && li < NOUT) {
        const int a = li / n_b, b = li - (li / n_b) * n_b;
        if (a < n_a) {
          const double* Ea = E_s[a];
          const double* Fv = F[b == 0 ? 0 : 1];
          const double Fm[9] = {Fv[0], Fv[3], Fv[4], Fv[3], Fv[1], Fv[5], Fv[4], Fv[5], Fv[2]};
          double M[9];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
              M[i * 3 + j] = Ea[i * 3] * Fm[j] + Ea[i * 3 + 1] * Fm[3 + j] + Ea[i * 3 + 2] * Fm[6 + j];
          double EF = 0.0;
#pragma unroll
          for (int t = 0; t < 9; ++t) EF = fma(Ea[t], Fm[t], EF);
          const double trE = Ea[0] + Ea[4] + Ea[8], trF = Fm[0] + Fm[4] + Fm[8];
          const double eps = src[S_EPS], e2 = eps * eps;
          const double* rb = pr.rb + ((int64_t)o * pr.na_tot + a) * 3;
          double rFn = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i)
            rFn += rb[i] * (Fm[i * 3] * nr[0] + Fm[i * 3 + 1] * nr[1] + Fm[i * 3 + 2] * nr[2]);
          T* c = &kd_s[grp][li * MHS_DC];
          c[0] = (T)Ea[0]; c[1] = (T)Ea[4]; c[2] = (T)Ea[8];
          c[3] = (T)Ea[1]; c[4] = (T)Ea[2]; c[5] = (T)Ea[5];
          c[6] = (T)Fv[0]; c[7] = (T)Fv[1]; c[8] = (T)Fv[2];
          c[9] = (T)Fv[3]; c[10] = (T)Fv[4]; c[11] = (T)Fv[5];
          c[12] = (T)M[0]; c[13] = (T)M[4]; c[14] = (T)M[8];
          c[15] = (T)(0.5 * (M[1] + M[3])); c[16] = (T)(0.5 * (M[2] + M[6])); c[17] = (T)(0.5 * (M[5] + M[7]));
          c[18] = (T)(pr.k3 * EF + trE * trF);
          c[19] = (T)(-3.0 * trF);
          c[20] = (T)(-3.0 * trE);
          c[21] = (T)(pr.k5e * e2 * EF);
          c[22] = (T)(pr.k7c * e2);
          c[23] = (T)((pr.flags & F_SUBTRACT) ? pr.k7e * (e2 * e2) * rFn : 0.0);
          c[24] = (T)pr.k5c;
        }
      }
    }
    __syncthreads();
    double acc[K][NOUT];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int t = 0; t < NOUT; ++t) acc[k][t] = 0.0;
    const T* kd = kd_s[grp];
    const T* kr = kr_s[grp];
    const T* kq = kq_s[grp];
    // 12 x 12 rules in whole chunks: NPT points per lane where LPP * NPT
    // divides 144, else chunks of 48 (16 lanes x 3 or 8 lanes x 6)
    constexpr int N12 = 144 % (LPP * NPT) == 0 ? NPT : LPP == 16 ? 3 : LPP == 8 ? 6 : NPT;
    // the direct series enters as -C1 times its sum
    if (valid && dfar) {
      if (nd == 12)
        far_pass<T, P, NOUT, N12, LPP, false, true, false, false>(pr, src, ob, 12, li, kd, kr, kq, -pr.C1, acc);
      else
        far_pass<T, P, NOUT, NPT, LPP, false, true, false, false>(pr, src, ob, nd, li, kd, kr, kq, -pr.C1, acc);
    }
    // the R and Q series run as separate passes even at equal orders: summed
    // in one pass their registers no longer fit the occupancy target
    if (valid && img && rfar) {
      if (ni == 12)
        far_pass<T, P, NOUT, N12, LPP, true, false, true, false>(pr, src, ob, 12, li, kd, kr, kq, 1.0, acc);
      else
        far_pass<T, P, NOUT, NPT, LPP, true, false, true, false>(pr, src, ob, ni, li, kd, kr, kq, 1.0, acc);
    }
    if (valid && img)
      far_pass<T, P, NOUT, NPT, LPP, true, false, false, true>(pr, src, ob, nq, li, kd, kr, kq, 1.0, acc);
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int t = 0; t < NOUT; ++t) {
        double r = acc[k][t];
        for (int off = LPP / 2; off > 0; off >>= 1) r += __shfl_xor(r, off, LPP);
        if (li == 0) res_s[(grp * K + k) * NOUT + t] = r;
      }
    __syncthreads();     // the kappas are rewritten by the next tile; res_s is complete
    {
      // the tile's PPB sources are PPB * K consecutive columns: one contiguous run of the row
      const int64_t s0 = ((int64_t)bx_ * pr.far_tiles + tile) * PPB;
      const int64_t oo = o - pr.out_lo;
      const int64_t n_here = (pr.n_src - s0 < PPB ? pr.n_src - s0 : PPB) * K * NOUT;
      const int os = pr.na_tot * pr.nb_tot;   // entries per (row, column); more than NOUT in a group launch
      double* dst = pr.out + (oo * pr.n_cols + s0 * K) * os;
      if (os == NOUT) {
        for (int64_t i = tid; i < n_here; i += 256) dst[i] = res_s[i];
      } else {
        for (int64_t i = tid; i < n_here; i += 256) dst[(i / NOUT) * os + i % NOUT] = res_s[i];
      }
    }
  }
}
template <typename T, int P, int NOUT, int NPT, int LPP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? MHS_FAR_WAVES32 : MHS_FAR_WAVES64)))
k_far(Params pr) {
  int64_t bx, by;
  rows_first(bx, by);
  body_far<T, P, NOUT, NPT, LPP>(pr, bx, by);
}
// the same pass over the tiles of a batch's queue (see drain_queue)
template <typename T, int P, int NOUT, int NPT, int LPP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? MHS_FAR_WAVES32 : MHS_FAR_WAVES64)))
kq_far(QueueArgs q) {
  drain_queue(q, [&](const Params& pr, int64_t bx, int64_t by) { body_far<T, P, NOUT, NPT, LPP>(pr, bx, by); });
}

// ------------------------------------------------------ mixed far pass --
// The mixed mode's far pass of the contracted interaction.  Its rule (Rule)
// gives far pairs a few Gauss points per family (one to a few dozen, against
// 64-144 in the FP64 rule), so a pair's point work no longer fills 8 lanes:
// a lane takes one (receiver row, source) pair and all of its points, in
// chunks of NPT.  The image kappas are not stored: each chunk forms
// kappa_k = sum_v F_v B[k][v] from the block's receiver tables in LDS (read
// by every lane at one address) as the series consumes it, so no per-pair
// table limits the occupancy.  Points, kappas and series in FP32; point
// offsets, weights and the sums over points in FP64.
#ifndef MHS_FARM_WAVES
#define MHS_FARM_WAVES 3
#endif
#ifndef MHS_FARM_NPT
#define MHS_FARM_NPT 4   // points per chunk of the mixed far pass
#endif
#ifndef MHS_FARM_NPT_MULTI
#define MHS_FARM_NPT_MULTI 2   // the same with two or more outputs (the fold's three slips; 1 and 4 ran no faster)
#endif

// kappa of output o, monomial k: sum_v F[o][v] B[o][8 k + v] (B a receiver
// table in LDS, rows of 8 floats of which 6 are used).
template <int M, int NOUT>
struct KapOnFly {
  const float* B[NOUT];
  float F[NOUT][6];
  __device__ __forceinline__ float operator[](int idx) const {
    const int o = idx / M, k = idx - (idx / M) * M;
    const float4 lo = *reinterpret_cast<const float4*>(B[o] + 8 * k);
    const float2 hi = *reinterpret_cast<const float2*>(B[o] + 8 * k + 4);
    return fmaf(F[o][5], hi.y, fmaf(F[o][4], hi.x, fmaf(F[o][3], lo.w,
           fmaf(F[o][2], lo.z, fmaf(F[o][1], lo.y, F[o][0] * lo.x)))));
  }
};

// One family of one pair over its ng x ng collapsed Gauss points, NPT at a
// time, summed times wscale into acc[k][t] for every node k (the series of
// a point is evaluated once for all K nodes).  FAM 0: direct series on the
// source triangle (kd: the per-output direct coefficients); 1: image R, 2:
// image Q on the mirrored triangle (kap: kappa on the fly).
template <int P, int NOUT, int NPT, int FAM, typename KapT>
__device__ __forceinline__ void mixed_pass(const Params& pr, const double* __restrict__ src, const double ob[3],
                                           int ng, const float* kd, const KapT& kap, double wscale,
                                           double (*acc)[NOUT]) {
  constexpr int K = (P + 1) * (P + 2) / 2;
  constexpr bool IMG = FAM != 0;
  const double* v = src + S_V;
  const double sz = IMG ? -1.0 : 1.0;
  const double ov[3] = {ob[0] - v[0], ob[1] - v[1], ob[2] - sz * v[2]};
  const double e1[3] = {v[3] - v[0], v[4] - v[1], sz * (v[5] - v[2])};
  const double e2[3] = {v[6] - v[0], v[7] - v[1], sz * (v[8] - v[2])};
  const float eps2 = (float)(src[S_EPS] * src[S_EPS]);
  const double ar2 = 2.0 * src[S_AREA] * wscale;
  const size_t toff = (size_t)(ng - 1) * ng * (2 * ng - 1) / 6;
  const double* qx1 = pr.gq + toff;
  const double* qx2 = qx1 + MHS_GQ_TOTAL;
  const double* qw = qx1 + 2 * MHS_GQ_TOTAL;
  const int npts = ng * ng;
  for (int base = 0; base < npts; base += NPT) {
    // compiler-only fence: the kappas are formed inside each chunk, where the
    // series consumes them, not hoisted out of the loop into registers
    __asm__ volatile("" ::: "memory");
    float x[NPT], y[NPT], z[NPT], ir[NPT], iq[NPT], sv[NOUT * NPT], b1[NPT], b2[NPT];
    double wn[NPT];
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
      const int g0 = base + u;
      const bool in = g0 < npts;
      const int g = in ? g0 : 0;
      const double x1 = qx1[g], x2 = qx2[g];
      wn[u] = in ? qw[g] * ar2 : 0.0;
      b1[u] = (float)x1;
      b2[u] = (float)x2;
      const float d1 = (float)fma(-x1, e1[0], fma(-x2, e2[0], ov[0]));
      const float d2 = (float)fma(-x1, e1[1], fma(-x2, e2[1], ov[1]));
      const float d3 = (float)fma(-x1, e1[2], fma(-x2, e2[2], ov[2]));
      x[u] = d1; y[u] = d2; z[u] = d3;
      radii(d1 * d1 + d2 * d2 + d3 * d3 + eps2, d3, FAM == 2, ir[u], iq[u]);
    }
    if constexpr (FAM == 0) d_blocked<float, NPT, NOUT>(kd, x, y, z, ir, sv);
    else if constexpr (FAM == 1) r_blocked_k<float, NPT, NOUT>(kap, x, y, z, ir, iq, sv);
    else q_blocked_k<float, NPT, NOUT>(kap, x, y, z, ir, iq, sv);
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
#pragma unroll
      for (int t = 0; t < NOUT; ++t) {
        if (P == 0) {
          acc[0][t] = fma(wn[u], (double)sv[t * NPT + u], acc[0][t]);
        } else {
          const double val = wn[u] * (double)sv[t * NPT + u];
          const float l1 = b1[u], l2 = b2[u], l0 = 1.0f - l1 - l2;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            float nk;
            if (P == 1) nk = k == 0 ? l0 : k == 1 ? l1 : l2;
            else if (k < 3) { const float l = k == 0 ? l0 : k == 1 ? l1 : l2; nk = l * (2.0f * l - 1.0f); }
            else nk = 4.0f * (k == 3 ? l0 * l1 : k == 4 ? l1 * l2 : l2 * l0);
            acc[k][t] = fma((double)nk, val, acc[k][t]);
          }
        }
      }
    }
  }
}

// Coefficients of the contracted direct series (MHS_DC floats, the layout of
// d_blocked) for a receiver half E and a source half F (symmetric, order 00
// 11 22 01 02 12), with the subtracted eigenstress term when eig (rFn = the
// receiver-frame row . F . receiver normal).
__device__ __forceinline__ void direct_series_coefs(const Params& pr, const double E[6], const double F[6],
                                                    double eps, double rFn, bool eig, float* c) {
  const double Em[9] = {E[0], E[3], E[4], E[3], E[1], E[5], E[4], E[5], E[2]};
  const double Fm[9] = {F[0], F[3], F[4], F[3], F[1], F[5], F[4], F[5], F[2]};
  double Mm[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Mm[i * 3 + j] = Em[i * 3] * Fm[j] + Em[i * 3 + 1] * Fm[3 + j] + Em[i * 3 + 2] * Fm[6 + j];
  double EF = 0.0;
#pragma unroll
  for (int q = 0; q < 9; ++q) EF = fma(Em[q], Fm[q], EF);
  const double trE = E[0] + E[1] + E[2], trF = F[0] + F[1] + F[2];
  const double e2 = eps * eps;
#pragma unroll
  for (int t = 0; t < 6; ++t) { c[t] = (float)E[t]; c[6 + t] = (float)F[t]; }
  c[12] = (float)Mm[0]; c[13] = (float)Mm[4]; c[14] = (float)Mm[8];
  c[15] = (float)(0.5 * (Mm[1] + Mm[3])); c[16] = (float)(0.5 * (Mm[2] + Mm[6]));
  c[17] = (float)(0.5 * (Mm[5] + Mm[7]));
  c[18] = (float)(pr.k3 * EF + trE * trF);
  c[19] = (float)(-3.0 * trF);
  c[20] = (float)(-3.0 * trE);
  c[21] = (float)(pr.k5e * e2 * EF);
  c[22] = (float)(pr.k7c * e2);
  c[23] = (float)(eig ? pr.k7e * (e2 * e2) * rFn : 0.0);
  c[24] = (float)pr.k5c;
}

// A family's points: one point alone (the order-1 rule) without the chunk's
// padding, else chunks of MHS_FARM_NPT.
template <int P, int NOUT, int FAM, typename KapT>
__device__ __forceinline__ void mixed_family(const Params& pr, const double* __restrict__ src, const double ob[3],
                                             int ng, const float* kd, const KapT& kap, double wscale,
                                             double (*acc)[NOUT]) {
  constexpr int NPT = NOUT == 1 ? MHS_FARM_NPT : MHS_FARM_NPT_MULTI;
  if (ng == 1) mixed_pass<P, NOUT, 1, FAM>(pr, src, ob, 1, kd, kap, wscale, acc);
  else mixed_pass<P, NOUT, NPT, FAM>(pr, src, ob, ng, kd, kap, wscale, acc);
}

template <int P, int NOUT>
__device__ __forceinline__ void body_far_mixed(const Params& pr, int64_t bx_, int64_t by_) {
  constexpr int K = (P + 1) * (P + 2) / 2;
  constexpr int KN = K * NOUT;
  __shared__ __align__(16) float BQ_s[2][MHS_QM * 8];
  __shared__ __align__(16) float BR_s[2][MHS_RM * 8];
  __shared__ double E_s[2][9];
  __shared__ double res_s[256 * KN];
  const int tid = threadIdx.x;
  const int n_a = pr.n_a, n_b = pr.n_b;
  const int64_t o = pr.obs_lo + by_;
  const double ob[3] = {pr.obs[3 * o], pr.obs[3 * o + 1], pr.obs[3 * o + 2]};
  const double* nr = pr.nrm + 3 * o;
  if (tid < n_a * 9) {
    const int a = tid / 9, i = (tid % 9) / 3, p = tid % 3;
    const double* rb = pr.rb + ((int64_t)o * pr.na_tot + a) * 3;
    const double rn = rb[0] * nr[0] + rb[1] * nr[1] + rb[2] * nr[2];
    E_s[a][i * 3 + p] = pr.mu * (rb[i] * nr[p] + rb[p] * nr[i]) + (i == p ? pr.lam * rn : 0.0);
  }
  __syncthreads();
  for (int t = tid; t < n_a * MHS_QM * 6; t += 256) {
    const int a = t / (MHS_QM * 6), kv = t % (MHS_QM * 6);
    double acc = 0.0;
    for (int r = mhs_qb_ptr[kv]; r < mhs_qb_ptr[kv + 1]; ++r)
      acc = fma(pr.coef[mhs_qb_coef[r]], E_s[a][mhs_qb_ip[r]], acc);
    BQ_s[a][(kv / 6) * 8 + kv % 6] = (float)acc;
  }
  {
    const double zp[3] = {1.0, ob[2], ob[2] * ob[2]};
    for (int t = tid; t < n_a * MHS_RM * 6; t += 256) {
      const int a = t / (MHS_RM * 6), kv = t % (MHS_RM * 6);
      double acc = 0.0;
      for (int r = mhs_rb_ptr[kv]; r < mhs_rb_ptr[kv + 1]; ++r)
        acc = fma(pr.coef[mhs_rb_coef[r]] * zp[mhs_rb_pz[r]], E_s[a][mhs_rb_ip[r]], acc);
      BR_s[a][(kv / 6) * 8 + kv % 6] = (float)acc;
    }
  }
  __syncthreads();
  for (int tile = 0; tile < pr.far_tiles; ++tile) {
    const int64_t s = ((int64_t)bx_ * pr.far_tiles + tile) * 256 + tid;
    double acc[K][NOUT];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int t = 0; t < NOUT; ++t) acc[k][t] = 0.0;
    if (s < pr.n_src) {
      const double* src = pr.src + s * MHS_SRC_STRIDE;
      const double eps = src[S_EPS];
      const double L = src[S_L];
      const double Dd = centroid_distance(src + S_C, ob, eps);
      const bool dfar = !(Dd <= pr.near_r * L);
      const bool img = !in_fold_band(src, ob, pr.flags);
      const double Di = centroid_distance(src + S_IC, ob, eps);
      const bool ifar = !(Di <= pr.near_r * src[S_IL]);
      // source half F_b (Hooke on the source plane, slip along frame row b)
      const double* n = src + S_N;
      double F[2][6];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (b < n_b) {
          const double* sb = pr.sb + (s * pr.nb_tot + b) * 3;
          half_tensor(sb, n, pr.mu, pr.lam, F[b]);
        } else {
#pragma unroll
          for (int v = 0; v < 6; ++v) F[b][v] = 0.0;
        }
      }
      int oa[NOUT], obb[NOUT];
#pragma unroll
      for (int t = 0; t < NOUT; ++t) { oa[t] = t / n_b; obb[t] = t - (t / n_b) * n_b; }
      // the image kappas' source halves, in FP32 (the FP64 ones die with the direct coefficients)
      KapOnFly<MHS_RM, NOUT> kr;
      KapOnFly<MHS_QM, NOUT> kq;
#pragma unroll
      for (int t = 0; t < NOUT; ++t) {
        const int a = oa[t] < n_a ? oa[t] : 0;
        kr.B[t] = BR_s[a];
        kq.B[t] = BQ_s[a];
#pragma unroll
        for (int v = 0; v < 6; ++v) kr.F[t][v] = kq.F[t][v] = (float)F[obb[t]][v];
      }
      if (dfar) {
        float kd[NOUT * MHS_DC];
#pragma unroll
        for (int t = 0; t < NOUT; ++t) {
          const double* Ea = E_s[oa[t] < n_a ? oa[t] : 0];
          const double* Fv = F[obb[t]];
          const double Fm[9] = {Fv[0], Fv[3], Fv[4], Fv[3], Fv[1], Fv[5], Fv[4], Fv[5], Fv[2]};
          double Mm[9];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
              Mm[i * 3 + j] = Ea[i * 3] * Fm[j] + Ea[i * 3 + 1] * Fm[3 + j] + Ea[i * 3 + 2] * Fm[6 + j];
          double EF = 0.0;
#pragma unroll
          for (int q = 0; q < 9; ++q) EF = fma(Ea[q], Fm[q], EF);
          const double trE = Ea[0] + Ea[4] + Ea[8], trF = Fm[0] + Fm[4] + Fm[8];
          const double e2 = eps * eps;
          const double* rb = pr.rb + ((int64_t)o * pr.na_tot + (oa[t] < n_a ? oa[t] : 0)) * 3;
          double rFn = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i)
            rFn += rb[i] * (Fm[i * 3] * nr[0] + Fm[i * 3 + 1] * nr[1] + Fm[i * 3 + 2] * nr[2]);
          float* c = kd + t * MHS_DC;
          c[0] = (float)Ea[0]; c[1] = (float)Ea[4]; c[2] = (float)Ea[8];
          c[3] = (float)Ea[1]; c[4] = (float)Ea[2]; c[5] = (float)Ea[5];
          c[6] = (float)Fv[0]; c[7] = (float)Fv[1]; c[8] = (float)Fv[2];
          c[9] = (float)Fv[3]; c[10] = (float)Fv[4]; c[11] = (float)Fv[5];
          c[12] = (float)Mm[0]; c[13] = (float)Mm[4]; c[14] = (float)Mm[8];
          c[15] = (float)(0.5 * (Mm[1] + Mm[3])); c[16] = (float)(0.5 * (Mm[2] + Mm[6]));
          c[17] = (float)(0.5 * (Mm[5] + Mm[7]));
          c[18] = (float)(pr.k3 * EF + trE * trF);
          c[19] = (float)(-3.0 * trF);
          c[20] = (float)(-3.0 * trE);
          c[21] = (float)(pr.k5e * e2 * EF);
          c[22] = (float)(pr.k7c * e2);
          c[23] = (float)((pr.flags & F_SUBTRACT) ? pr.k7e * (e2 * e2) * rFn : 0.0);
          c[24] = (float)pr.k5c;
        }
        const int nd = rule_order(pr.rule, Dd, L) + P;
        mixed_family<P, NOUT, 0>(pr, src, ob, nd, kd, 0, -pr.C1, acc);
      }
      if (img) {
        if (ifar) {
          const int ni = rule_order(pr.rule, Di, src[S_IL]) + P;
          mixed_family<P, NOUT, 1>(pr, src, ob, ni, nullptr, kr, 1.0, acc);
        }
        // the Q-family has no closed form: near pairs keep the budget law
        const int nq = ifar ? rule_order(pr.rule, Di, src[S_IL]) + P : q_order_fast<P>(src, ob);
        mixed_family<P, NOUT, 2>(pr, src, ob, nq, nullptr, kq, 1.0, acc);
      }
    }
    const int64_t oo = o - pr.out_lo;
    const int os = pr.na_tot * pr.nb_tot;   // entries per (row, column); more than NOUT in a group launch
    if (KN == 1 && os == 1) {
      if (s < pr.n_src) pr.out[oo * pr.n_cols + s] = acc[0][0];
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int t = 0; t < NOUT; ++t) res_s[(tid * K + k) * NOUT + t] = acc[k][t];
      __syncthreads();
      // the tile's sources are 256 K consecutive columns: one contiguous run of the row
      const int64_t s0 = ((int64_t)bx_ * pr.far_tiles + tile) * 256;
      const int64_t n_here = (pr.n_src - s0 < 256 ? pr.n_src - s0 : 256) * KN;
      double* dst = pr.out + (oo * pr.n_cols + s0 * K) * os;
      if (os == NOUT) {
        for (int64_t i = tid; i < n_here; i += 256) dst[i] = res_s[i];
      } else {
        for (int64_t i = tid; i < n_here; i += 256) dst[(i / NOUT) * os + i % NOUT] = res_s[i];
      }
      __syncthreads();
    }
  }
}
template <int P, int NOUT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MHS_FARM_WAVES)))
k_far_mixed(Params pr) {
  int64_t bx, by;
  rows_first(bx, by);
  body_far_mixed<P, NOUT>(pr, bx, by);
}
// the same pass over the tiles of a batch's queue (see drain_queue)
template <int P, int NOUT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MHS_FARM_WAVES)))
kq_far_mixed(QueueArgs q) {
  drain_queue(q, [&](const Params& pr, int64_t bx, int64_t by) { body_far_mixed<P, NOUT>(pr, bx, by); });
}

// Near parts of the contracted interaction (closed forms, folded mirror),
// added after k_far.

// One pair's near parts as a real call: inlined into the list loop its
// closed forms spill far more than as a function of its own.
template <int P, int NM>
__device__ __attribute__((noinline)) void near_item_call(const Params& pr, int64_t o, int64_t col) {
  pair_near_item<P, NM>(pr, o, col);
}

// A block is one receiver row x MHS_ROW_TILE node columns; about one pair in
// nine has a near part, so the block lists its columns that do (those with
// an image or mirror part first, then the direct-only ones) and evaluates the
// list with every lane busy, a column per thread.  (Evaluating all K node
// columns of a source from one set of closed-form raw moments held in scratch
// was faster at order 2 while the image edge integrals ran as a memory-bound
// loop, and slower once they were unrolled: 84 against 66 ms at n = 1560.)
template <int P, int NM, int NT>
__device__ __forceinline__ void body_near(const Params& pr, int64_t bx_, int64_t by_) {
  static_assert(NT == 64 || NT == 256, "a near block is one wave or four");
  constexpr int K = (P + 1) * (P + 2) / 2;
  __shared__ int list_s[2 * MHS_ROW_TILE];
  const int64_t o = pr.obs_lo + by_;
  const int64_t c0 = (int64_t)bx_ * MHS_ROW_TILE;
  const double ob[3] = {pr.obs[3 * o], pr.obs[3 * o + 1], pr.obs[3 * o + 2]};
  int n_img, n_dir;
  // a step of 256 columns whose sources' centroids and mirrored centroids
  // all lie farther than the near switch (none straddling) has no near part
  auto skip = [&](int64_t col0) {
    if (!pr.near_tiles) return false;
    const double* t = pr.near_tiles + 8 * (col0 / 256);
    if (t[7] != 0.0) return false;
    const double dx = fmax(fmax(t[0] - ob[0], ob[0] - t[3]), 0.0);
    const double dy = fmax(fmax(t[1] - ob[1], ob[1] - t[4]), 0.0);
    const double dz = fmax(fmax(t[2] - ob[2], ob[2] - t[5]), 0.0);
    return sqrt(dx * dx + dy * dy + dz * dz) > pr.near_r * t[6] * (1.0 + 1e-9);
  };
  auto cls = [&](int64_t col) {
    const int parts = near_parts(pr, pr.src + (col / K) * MHS_SRC_STRIDE, ob);
    return (parts & 2) ? 1 : parts == 1 ? 2 : 0;
  };
  if constexpr (NT == 64) {
    list_row_tile_wave(pr.n_cols, c0, cls, list_s, n_img, n_dir, skip);
    __syncthreads();
  } else {
    list_row_tile(pr.n_cols, c0, cls, list_s, n_img, n_dir, skip);
  }
  for (int i = threadIdx.x; i < n_img + n_dir; i += NT) {
    __asm__ volatile("" ::: "memory");   // no loads hoisted out of the pair loop
    const int c = i < n_img ? list_s[i] : list_s[MHS_ROW_TILE + i - n_img];
    near_item_call<P, NM>(pr, o, c0 + c);
  }
}
template <int P, int NM>
__global__ void __launch_bounds__(MHS_NEAR_THREADS) __attribute__((amdgpu_waves_per_eu(MHS_NEAR_WAVES)))
k_near(Params pr) {
  int64_t bx, by;
  rows_first(bx, by);
  body_near<P, NM, MHS_NEAR_THREADS>(pr, bx, by);
}
// the same pass over the tiles of a batch's queue (see drain_queue)
template <int P, int NM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MHS_NEAR_WAVES)))
kq_near(QueueArgs q) {
  drain_queue(q, [&](const Params& pr, int64_t bx, int64_t by) { body_near<P, NM, 256>(pr, bx, by); });
}

__global__ void k_zero_i32(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Append every (observer row, straddling source) pair inside the fold band.
__global__ void k_fold_list(Params pr, const int64_t* __restrict__ strad, int64_t n_strad,
                            int64_t n_rows, int64_t* __restrict__ list, int* __restrict__ count,
                            int64_t cap) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_rows * n_strad; t += stride) {
    const int64_t o = pr.obs_lo + t / n_strad;
    const int64_t s = strad[t % n_strad];
    const double* src = pr.src + s * MHS_SRC_STRIDE;
    const double ob[3] = {pr.obs[3 * o], pr.obs[3 * o + 1], pr.obs[3 * o + 2]};
    if (!in_fold_band(src, ob, pr.flags)) continue;
    const int slot = atomicAdd(count, 1);
    if (slot < cap) {
      list[2 * slot] = o;
      list[2 * slot + 1] = s;
    }
  }
}

// Prepare every folded pair in parallel, a thread each: its polar grid and
// the surface load at the receiver's foot per node, into the work buffer.
template <int P>
__global__ void __launch_bounds__(64) k_fold_prep(Params pr, const int64_t* __restrict__ list,
                                                  const int* __restrict__ count, int64_t cap,
                                                  FoldItem* __restrict__ work, int64_t work_items,
                                                  int* __restrict__ err) {
  constexpr int K = (P + 1) * (P + 2) / 2;
  int64_t n_items = *count;
  if (n_items > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, 2);
    n_items = cap;
  }
  if (n_items > work_items) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, 4);
    n_items = work_items;
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; item < n_items; item += stride) {
    const int64_t o = list[2 * item], s = list[2 * item + 1];
    const double* src = pr.src + s * MHS_SRC_STRIDE;
    const double xh0 = pr.obs[3 * o], xh1 = pr.obs[3 * o + 1];
    const double zeta = fmax(0.0, -pr.obs[3 * o + 2]);
    FoldItem& it = work[item];
    fold_grid_build(it.G, xh0, xh1, zeta, src[S_CREASE], src[S_CREASE + 1], src[S_CREASE_R], src[S_EPS],
                    src[S_DIAM]);
    if (it.G.overflow) atomicOr(err, 1);
    FoldContract fc;
    {
      Frame f;
      load_frame(src, false, f);
      fold_contract_prep(pr, f, src, fc);
    }
    for (int k = 0; k < K; ++k) surface_load<P>(pr, src, fc, k, xh0, xh1, it.s0[k]);
  }
}

// The closed-form surface load as a function of its own: inlined into the
// point loop it spills more and runs slower.
template <int P>
__device__ __attribute__((noinline)) void surface_load_call(const Params& pr, const double* __restrict__ src,
                                                            const FoldContract& fc, int k, double x, double y,
                                                            double* sl) {
  surface_load<P>(pr, src, fc, k, x, y, sl);
}

// The Boussinesq sum of the folded pairs over their polar grids, split by
// point: NEAR the points within the closed-form range of the source
// (sqrt(|x - c|^2 + eps^2) <= 10 L) plus Love's disk term, FAR the others by
// the Gauss tables only (a kernel without the closed form's registers).
// Each adds its part; together they are the whole sum.  A block first
// compacts the indices of its kind of point into an LDS list (in ascending
// grid order, so the summation order is fixed), then evaluates the list with
// every lane busy: at a surface-reaching fault about 92 % of the points are
// near, and without the list most far-kernel lanes would idle.
// With the mixed rule (MIX, far points only) the near/far switch is the
// rule's and a far point's surface load is the contracted direct series at
// the rule's order in FP32 points (sums in FP64), per slip direction.
#define MHS_FOLD_LIST 4096
template <int P, bool STRESS, bool NEAR, bool MIX = false>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(NEAR ? MHS_FOLD_NEAR_WAVES : MHS_FOLD_FAR_WAVES)))
k_fold(Params pr, const int64_t* __restrict__ list,
                                              const int* __restrict__ count, int64_t cap,
                                              const FoldItem* __restrict__ work, int64_t work_items) {
  static_assert(!(NEAR && MIX), "the mixed rule changes the far points only");
  constexpr int K = (P + 1) * (P + 2) / 2;
  constexpr int NA = STRESS ? 18 : 9;
  constexpr int GW = sizeof(FoldGrid) / 8;
  static_assert(sizeof(FoldGrid) % 8 == 0, "FoldGrid copies as 8-byte words");
  __shared__ FoldGrid G;
  __shared__ float kdz_s[MIX ? 3 * MHS_DC : 1];   // the item's direct-series coefficients, per slip
  __shared__ FoldContract fc_s;                    // the item's surface-load contraction (not MIX)
  __shared__ double sh[3];
  __shared__ double red[NA][8];
  __shared__ double u6[6], w6[6], u4[4], w4[4];
  __shared__ int list_s[MHS_FOLD_LIST];
  __shared__ int wcnt_s[4];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  int64_t n_items = *count;
  if (n_items > cap) n_items = cap;
  if (n_items > work_items) n_items = work_items;
  if (tid < 6) { u6[tid] = pr.gl_x[15 + tid]; w6[tid] = pr.gl_w[15 + tid]; }
  if (tid < 4) { u4[tid] = pr.gl_x[6 + tid]; w4[tid] = pr.gl_w[6 + tid]; }
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int64_t o = list[2 * item], s = list[2 * item + 1];
    const double* src = pr.src + s * MHS_SRC_STRIDE;
    const double xh0 = pr.obs[3 * o], xh1 = pr.obs[3 * o + 1];
    const double zeta = fmax(0.0, -pr.obs[3 * o + 2]);
    __syncthreads();
    {
      const unsigned long long* gs = reinterpret_cast<const unsigned long long*>(&work[item].G);
      unsigned long long* gd = reinterpret_cast<unsigned long long*>(&G);
      for (int w = tid; w < GW; w += blockDim.x) gd[w] = gs[w];
    }
    if (MIX && tid < 3) {
      // sigma_zz on the plane z = 0 per unit slip e_tid, minus the eigenstress (fold.py:214-228)
      const double ez[3] = {0.0, 0.0, 1.0};
      const double em[3] = {tid == 0 ? 1.0 : 0.0, tid == 1 ? 1.0 : 0.0, tid == 2 ? 1.0 : 0.0};
      double E[6], Fm[6];
      half_tensor(ez, ez, pr.mu, pr.lam, E);
      half_tensor(em, src + S_N, pr.mu, pr.lam, Fm);
      direct_series_coefs(pr, E, Fm, src[S_EPS], Fm[2], true, kdz_s + tid * MHS_DC);
    }
    if (!MIX && tid == 0) {
      Frame f;
      load_frame(src, false, f);
      fold_contract_prep(pr, f, src, fc_s);
    }
    __syncthreads();
    for (int k = 0; k < K; ++k) {
      if (tid < 3) sh[tid] = work[item].s0[k][tid];
      __syncthreads();
      double acc[NA];
#pragma unroll
      for (int t = 0; t < NA; ++t) acc[t] = 0.0;
      const double s0 = sh[0], s1 = sh[1], s2 = sh[2];
      const int total = G.total;
      int g0 = 0;
      while (g0 < total) {
        // compaction: whole steps of 256 points while the list has room
        int n = 0;
        while (g0 < total && n <= MHS_FOLD_LIST - 256) {
          const int g = g0 + tid;
          bool take = false;
          if (g < total) {
            double x, y, w;
            fold_grid_point(G, xh0, xh1, u6, w6, u4, w4, g, x, y, w);
            const double op[3] = {x, y, 0.0};
            take = (centroid_distance(src + S_C, op, src[S_EPS]) <= pr.fold_near_r * src[S_L]) == NEAR;
          }
          const unsigned long long m = __ballot(take);
          if ((tid & 63) == 0) wcnt_s[wave] = __popcll(m);
          __syncthreads();
          int off = n, step = 0;
#pragma unroll
          for (int wv = 0; wv < 4; ++wv) {
            off += wv < wave ? wcnt_s[wv] : 0;
            step += wcnt_s[wv];
          }
          if (take)
            list_s[off + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = g;
          n += step;
          g0 += 256;
          __syncthreads();        // wcnt_s is rewritten by the next step
        }
        for (int i = tid; i < n; i += 256) {
          __asm__ volatile("" ::: "memory");   // no loads hoisted out of the point loop
          const int g = list_s[i];
          double x, y, w;
          fold_grid_point(G, xh0, xh1, u6, w6, u4, w4, g, x, y, w);
          double sl[3];
          if constexpr (NEAR) {
            surface_load_call<P>(pr, src, fc_s, k, x, y, sl);
          } else if constexpr (MIX) {
            const double op[3] = {x, y, 0.0};
            const int ng = rule_order(pr.rule, centroid_distance(src + S_C, op, src[S_EPS]), src[S_L]) + P;
            double am[K][3];
#pragma unroll
            for (int kk = 0; kk < K; ++kk) am[kk][0] = am[kk][1] = am[kk][2] = 0.0;
            mixed_family<P, 3, 0>(pr, src, op, ng, kdz_s, 0, -2.0 * pr.C1, am);
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
              if (kk == k) { sl[0] = am[kk][0]; sl[1] = am[kk][1]; sl[2] = am[kk][2]; }
          } else {
            surface_load_far<P>(pr, src, fc_s, k, x, y, sl);
          }
          sl[0] -= s0; sl[1] -= s1; sl[2] -= s2;
          const double dx = xh0 - x, dy = xh1 - y;
          if (STRESS) {
            double S[6];
            point_stress(dx, dy, zeta, pr.nu, S);
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
              for (int m = 0; m < 3; ++m) acc[c * 3 + m] = fma(w * S[c], sl[m], acc[c * 3 + m]);
          } else {
            double U[3];
            point_disp(dx, dy, zeta, pr.mu, pr.nu, U);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
              for (int m = 0; m < 3; ++m) acc[c * 3 + m] = fma(w * U[c], sl[m], acc[c * 3 + m]);
          }
        }
        __syncthreads();          // the list is refilled by the next round
      }
      // block reduction: wave shuffle, then across the 4 waves
#pragma unroll
      for (int t = 0; t < NA; ++t) {
        double v = acc[t];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((tid & 63) == 0) red[t][tid >> 6] = v;
      }
      __syncthreads();
      if (tid == 0) {
        const int nw = (blockDim.x + 63) / 64;
        double tot[NA];
        for (int t = 0; t < NA; ++t) {
          double v = 0.0;
          for (int wv = 0; wv < nw; ++wv) v += red[t][wv];
          tot[t] = v;
        }
        const int64_t col = s * K + k;
        if (STRESS) {
          double srr = 0.0, szz = 0.0;
          if (NEAR) disk_stress(G.R1, zeta, pr.nu, srr, szz);
          // components (xx, yy, xy, xz, yz, zz) -> H[a][b][m]
          double H[27];
          const int ci[3][3] = {{0, 2, 3}, {2, 1, 4}, {3, 4, 5}};
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              for (int m = 0; m < 3; ++m) {
                double disk = (a == b) ? (a == 2 ? szz : srr) : 0.0;
                H[(a * 3 + b) * 3 + m] = tot[ci[a][b] * 3 + m] + sh[m] * disk;
              }
          emit_stress(pr, o, col, s, H, 0.0, 0.0, src + S_N, src + S_N, true);
        } else {
          const double uz = NEAR ? disk_uz(G.R1, zeta, pr.mu, pr.nu) : 0.0;
          double U[9];
          for (int i = 0; i < 3; ++i)
            for (int m = 0; m < 3; ++m) U[i * 3 + m] = tot[i * 3 + m] + (i == 2 ? sh[m] * uz : 0.0);
          emit_disp(pr, o, col, U, true);
        }
      }
      __syncthreads();
    }
  }
}


// Point counts of the fold's passes (the work model of their roofline), from
// the pair list and polar grids a launch left in its fold buffers: per
// folded (pair, node column) the points the near kernel evaluates in closed
// form and the far kernel by Gauss tables, and the far points' Gauss points
// (the rule's order squared), by the kernels' own classification.
// counts[0] near points, [1] far points, [2] far Gauss points, [3] folded
// (pair, node) items; added into.
template <int P>
__global__ void __launch_bounds__(256) k_fold_hist(Params pr, const int64_t* __restrict__ list,
                                                   const int* __restrict__ count, int64_t cap,
                                                   const FoldItem* __restrict__ work, int64_t work_items,
                                                   unsigned long long* __restrict__ counts) {
  constexpr int K = (P + 1) * (P + 2) / 2;
  constexpr int GW = sizeof(FoldGrid) / 8;
  __shared__ FoldGrid G;
  __shared__ double u6[6], w6[6], u4[4], w4[4];
  const int tid = threadIdx.x;
  int64_t n_items = *count;
  if (n_items > cap) n_items = cap;
  if (n_items > work_items) n_items = work_items;
  if (tid < 6) { u6[tid] = pr.gl_x[15 + tid]; w6[tid] = pr.gl_w[15 + tid]; }
  if (tid < 4) { u4[tid] = pr.gl_x[6 + tid]; w4[tid] = pr.gl_w[6 + tid]; }
  unsigned long long nn = 0, nf = 0, ng2 = 0;
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int64_t o = list[2 * item], s = list[2 * item + 1];
    const double* src = pr.src + s * MHS_SRC_STRIDE;
    const double xh0 = pr.obs[3 * o], xh1 = pr.obs[3 * o + 1];
    __syncthreads();
    {
      const unsigned long long* gs = reinterpret_cast<const unsigned long long*>(&work[item].G);
      unsigned long long* gd = reinterpret_cast<unsigned long long*>(&G);
      for (int w = tid; w < GW; w += blockDim.x) gd[w] = gs[w];
    }
    __syncthreads();
    for (int g = tid; g < G.total; g += 256) {
      double x, y, w;
      fold_grid_point(G, xh0, xh1, u6, w6, u4, w4, g, x, y, w);
      const double op[3] = {x, y, 0.0};
      const double D = centroid_distance(src + S_C, op, src[S_EPS]);
      if (D <= pr.fold_near_r * src[S_L]) {
        ++nn;
      } else {
        ++nf;
        const int ng = pr.mixed_rule ? rule_order(pr.rule, D, src[S_L]) + P : (D > 40.0 * src[S_L] ? 8 : 12);
        ng2 += (unsigned long long)(ng * ng);
      }
    }
  }
  if (nn) atomicAdd(counts + 0, nn * K);
  if (nf) atomicAdd(counts + 1, nf * K);
  if (ng2) atomicAdd(counts + 2, ng2 * K);
  if (tid == 0 && blockIdx.x == 0) atomicAdd(counts + 3, (unsigned long long)n_items * K);
}

// ------------------------------------------------------------- histogram --
// Branch histogram of the pair set, with the kernels' own classification
// (the work model behind the roofline figures).  counts[0..12], the FP64
// rule: pairs, direct near / far12 / far8, image near / far12 / far8, fold
// pairs, Q points, Q pairs, mirror near / far12 / far8.  counts[13..23], the
// mixed rule of the contracted interaction (zero otherwise), per (row,
// source) pair: pairs, direct near / far pairs / far points, image R near /
// far pairs / far points, Q pairs / points / chunks of the mixed far pass,
// R chunks.
#define MHS_HIST_N 24
__global__ void k_hist(Params pr, int order, unsigned long long* __restrict__ counts) {
  __shared__ unsigned long long acc[MHS_HIST_N];
  if (threadIdx.x < MHS_HIST_N) acc[threadIdx.x] = 0ull;
  __syncthreads();
  const int K = (order + 1) * (order + 2) / 2;
  const int64_t o = pr.obs_lo + blockIdx.y;
  const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (col < pr.n_cols) {
    const int64_t s = col / K;
    const double* src = pr.src + s * MHS_SRC_STRIDE;
    const double ob[3] = {pr.obs[3 * o], pr.obs[3 * o + 1], pr.obs[3 * o + 2]};
    const bool band = in_fold_band(src, ob, pr.flags);
    const double eps = src[S_EPS], L = src[S_L];
    atomicAdd(&acc[0], 1ull);
    for (int img = 0; img < 2; ++img) {
      const double D = centroid_distance(src + (img ? S_IC : S_C), ob, eps);
      const int cls = D <= 10.0 * L ? 0 : (D > 40.0 * L ? 2 : 1);
      const int base = img ? (band ? 10 : 4) : 1;
      atomicAdd(&acc[base + cls], 1ull);
    }
    if (band) {
      atomicAdd(&acc[7], 1ull);
    } else {
      const int nq = order == 0 ? q_order_fast<0>(src, ob)
                                : (order == 1 ? q_order_fast<1>(src, ob) : q_order_fast<2>(src, ob));
      atomicAdd(&acc[8], (unsigned long long)(nq * nq));
      atomicAdd(&acc[9], 1ull);
    }
    const bool mixed_rule = (pr.flags & F_MIXED) && pr.kind == K_INTERACTION && pr.sb && pr.n_a * pr.n_b <= 2;
    if (mixed_rule && col == s * K) {
      // the mixed far pass's branches (k_far_mixed), once per (row, source) pair
      constexpr int NPT = MHS_FARM_NPT;
      const int nptf = pr.n_a * pr.n_b == 1 ? NPT : (NPT + 1) / 2;
      auto chunks = [&](int n) { return n == 1 ? 1 : (n * n + nptf - 1) / nptf; };
      atomicAdd(&acc[13], 1ull);
      const double Dd = centroid_distance(src + S_C, ob, eps);
      if (Dd <= pr.near_r * L) {
        atomicAdd(&acc[14], 1ull);
      } else {
        const int nd = rule_order(pr.rule, Dd, L) + order;
        atomicAdd(&acc[15], 1ull);
        atomicAdd(&acc[16], (unsigned long long)(nd * nd));
      }
      if (!band) {
        const double Di = centroid_distance(src + S_IC, ob, eps);
        const bool ifar = !(Di <= pr.near_r * src[S_IL]);
        int nq;
        if (ifar) {
          const int ni = rule_order(pr.rule, Di, src[S_IL]) + order;
          atomicAdd(&acc[18], 1ull);
          atomicAdd(&acc[19], (unsigned long long)(ni * ni));
          atomicAdd(&acc[23], (unsigned long long)chunks(ni));
          nq = ni;
        } else {
          atomicAdd(&acc[17], 1ull);
          nq = order == 0 ? q_order_fast<0>(src, ob)
                          : (order == 1 ? q_order_fast<1>(src, ob) : q_order_fast<2>(src, ob));
        }
        atomicAdd(&acc[20], 1ull);
        atomicAdd(&acc[21], (unsigned long long)(nq * nq));
        atomicAdd(&acc[22], (unsigned long long)chunks(nq));
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < MHS_HIST_N) atomicAdd(&counts[threadIdx.x], acc[threadIdx.x]);
}

// ---------------------------------------------------------------- gather --
// Continuous layout: out[..., g, ...] = sum over the element columns c of
// shared node g (CSR, ascending) of scratch[..., c, ...].  Interaction
// additionally resolves the Cartesian slip on node g's own frame.
__global__ void k_gather(const double* __restrict__ scr, double* __restrict__ out, int64_t n_lead,
                         int64_t n_elem_cols, int64_t n_nodes, int tail, const int64_t* __restrict__ ptr,
                         const int64_t* __restrict__ idx, const double* __restrict__ sb, int n_a, int n_b) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (sb == nullptr) {
    // layout (n_lead, n_cols, tail)
    const int64_t total = n_lead * n_nodes * tail;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const int64_t e = t % tail;
    const int64_t g = (t / tail) % n_nodes;
    const int64_t L = t / (tail * n_nodes);
    double acc = 0.0;
    for (int64_t q = ptr[g]; q < ptr[g + 1]; ++q) acc += scr[(L * n_elem_cols + idx[q]) * tail + e];
    out[t] = acc;
    }
  } else {
    // interaction: scratch (n_obs, n_elem_cols, n_a, 3) -> out (n_obs, n_nodes, n_a, n_b)
    const int64_t total = n_lead * n_nodes * n_a * n_b;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const int b = (int)(t % n_b);
    const int a = (int)((t / n_b) % n_a);
    const int64_t g = (t / ((int64_t)n_a * n_b)) % n_nodes;
    const int64_t L = t / ((int64_t)n_a * n_b * n_nodes);
    const double* bv = sb + (g * n_b + b) * 3;
    double acc = 0.0;
    for (int64_t q = ptr[g]; q < ptr[g + 1]; ++q) {
      const double* v = scr + ((L * n_elem_cols + idx[q]) * n_a + a) * 3;
      acc += v[0] * bv[0] + v[1] * bv[1] + v[2] * bv[2];
    }
    out[t] = acc;
    }
  }
}

// Elastic strain from stress (matrices.py:572-579): tr(e) = tr(sig)/(3 lam + 2 mu),
// e = (sig - lam tr(e) I) / (2 mu), in place on (n_obs, 3, 3, n_cols, 3).
// Source-major copy of an interaction matrix: out[s][b][a][r] = in[r][s][a][b]
// (in n_rec x n_src x n_a x n_b, out n_src x n_b x n_a x n_rec).  A block moves
// a tile of MHS_TR_R receivers x ts = MHS_TR_LD / (n_a n_b) sources through
// LDS: each receiver's ts * n_a * n_b entries are one contiguous read, each
// (source, b, a) row of the tile's receivers one contiguous write.  HBM-bound
// (8 bytes read and 8 written per entry).
#define MHS_TR_R 32
#define MHS_TR_LD 64
__global__ void __launch_bounds__(256) k_source_major(const double* __restrict__ in, double* __restrict__ out,
                                                      int64_t n_rec, int64_t n_src, int n_a, int n_b) {
  __shared__ double tile[MHS_TR_R][MHS_TR_LD + 1];
  const int nab = n_a * n_b, ts = MHS_TR_LD / nab;
  const int64_t r0 = (int64_t)blockIdx.y * MHS_TR_R, s0 = (int64_t)blockIdx.x * ts;
  const int rows = (int)(n_rec - r0 < MHS_TR_R ? n_rec - r0 : MHS_TR_R);
  const int cols = (int)(n_src - s0 < ts ? n_src - s0 : ts);
  const int w = cols * nab;                     // contiguous entries per receiver in the tile
  for (int i = threadIdx.x; i < rows * MHS_TR_LD; i += 256) {
    const int r = i / MHS_TR_LD, c = i - r * MHS_TR_LD;
    if (c < w) tile[r][c] = in[((r0 + r) * n_src + s0) * nab + c];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < cols * nab * MHS_TR_R; i += 256) {
    const int q = i / MHS_TR_R, r = i - q * MHS_TR_R;   // q = (source, b, a) within the tile
    if (r < rows) {
      const int s = q / nab, ba = q - s * nab, b = ba / n_a, a = ba - b * n_a;
      out[((s0 + s) * nab + ba) * n_rec + r0 + r] = tile[r][s * nab + a * n_b + b];
    }
  }
}

// a -= b elementwise (the elastic stress from the finished total and
// eigenstress).
__global__ void k_sub(double* __restrict__ a, const double* __restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) a[t] = a[t] - b[t];
}

__global__ void k_strain(double* __restrict__ buf, int64_t n_obs, int64_t n_cols, double mu, double lam) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_obs * n_cols * 3; t += stride) {
    const int64_t j = t % 3;
    const int64_t c = (t / 3) % n_cols;
    const int64_t o = t / (3 * n_cols);
    auto at = [&](int a, int b) -> double& { return buf[((o * 9 + a * 3 + b) * n_cols + c) * 3 + j]; };
    const double tre = (at(0, 0) + at(1, 1) + at(2, 2)) / (3.0 * lam + 2.0 * mu);
    for (int a = 0; a < 3; ++a) at(a, a) -= lam * tre;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) at(a, b) /= (2.0 * mu);
  }
}

// The mixed mode's default far-field rule: closed forms up to 4 L, then 5 x 5
// Gauss points per family up to 6 L, 4 x 4 up to 16 L, 3 x 3 up to 100 L and
// 2 x 2 beyond (plus the nodal order per direction).  Every pair's
// quadrature error stays below 1e-7 of the pair's own largest entry, for
// every form and order (worst measured: the traction of a constant-slip
// source against 24 x 24 Gauss, 6e-8 at 100 L), so a row's error is bounded
// relative to the row wherever its observer lies.
#ifndef MHS_FARM_SINGLE
#define MHS_FARM_SINGLE 1   // the mixed far pass in one-output groups
#endif
#ifndef MHS_FARM_CT
#define MHS_FARM_CT 2   // default column tiles (of 256 sources) per block of the mixed far pass
#endif
static Rule default_rule() {
  Rule r;
  r.near = 4.0;
  const double hi[MHS_RULE_BANDS] = {6.0, 16.0, 100.0, 1e300, 1e300, 1e300, 1e300, 1e300};
  const int n[MHS_RULE_BANDS] = {5, 4, 3, 2, 2, 2, 2, 2};
  for (int b = 0; b < MHS_RULE_BANDS; ++b) { r.hi[b] = hi[b]; r.n[b] = n[b]; }
  return r;
}

// Whether an interaction launch runs the contracted passes (pointwise far
// field per output group of one or two components, each point once for all
// K node columns) or the tensor passes (moment tables lifted to all 27
// stress components, per node column).  One or two components: contracted.
// More, in the mixed mode: contracted while its output groups (a receiver
// row with up to two source rows each) are no more than the node columns K
// -- measured: the three-fault network at order 2 (3 x 2 components, K = 6)
// 1.03 s -> 0.55 s, the Gaussian-field fault at order 0 (3 x 3, K = 1)
// 10 ms -> 16 ms.  With the far pass in one-output launches the rule still
// holds: every mixed 3 x 3 interaction contracted ran the Gaussian-field
// fault at order 0 in 20.8 ms against 11.4 ms, and a 1,440-triangle fault
// in 17.7 / 69.2 / 229 ms against 8.3 / 59.7 / 227 ms at orders 0 / 1 / 2.
// The FP64 mode keeps the tensor passes beyond two components (faster
// there: 2.2 s against 2.8 s on the network).
static bool use_contracted(int kind, bool has_sb, int n_a, int n_b, int order, int flags) {
  if (kind != K_INTERACTION || !has_sb) return false;
  if (n_a * n_b <= 2) return true;
  const int K = (order + 1) * (order + 2) / 2;
  const int groups = n_a * ((n_b + 1) / 2);
  return (flags & F_MIXED) && groups <= K;
}

// The near/far switch and the far-field rule of a launch: the reference's
// 10 L, or with MHS_MIXED on the contracted interaction the descriptor's
// rule (checked) or the default.
static int fill_rule(const mhs_launch* L, Params& pr) {
  const bool mixed_rule = (L->flags & F_MIXED) != 0;
  pr.near_r = 10.0;
  pr.fold_near_r = 10.0;
  pr.rule = default_rule();
  pr.mixed_rule = mixed_rule ? 1 : 0;
  if (!mixed_rule) return 0;
  if (L->rule.near_ratio != 0.0) {
    const mhs_rule& r = L->rule;
    if (!(r.near_ratio >= 1.0 && r.near_ratio <= 10.0))
      return fail("mhs_assemble: rule.near_ratio must be in [1, 10] (or 0 for the default rule)");
    for (int b = 0; b < MHS_RULE_BANDS; ++b) {
      if (r.order[b] < 1 || r.order[b] > 190) return fail("mhs_assemble: rule orders must be in 1..190");
      if (b > 0 && !(r.band_hi[b] >= r.band_hi[b - 1])) return fail("mhs_assemble: rule bands must ascend");
      pr.rule.hi[b] = r.band_hi[b];
      pr.rule.n[b] = r.order[b];
    }
    pr.rule.near = r.near_ratio;
    if (r.fold_near_ratio != 0.0 && !(r.fold_near_ratio >= 1.0 && r.fold_near_ratio <= 10.0))
      return fail("mhs_assemble: rule.fold_near_ratio must be in [1, 10] (or 0 for near_ratio)");
  }
  pr.near_r = pr.rule.near;
  pr.fold_near_r = (L->rule.near_ratio != 0.0 && L->rule.fold_near_ratio != 0.0) ? L->rule.fold_near_ratio
                                                                                 : pr.near_r;
  return 0;
}

// Row-chunked 2-D launches: blockIdx.y is the observer row within a chunk of
// at most MHS_ROWS_PER_LAUNCH rows and blockIdx.x the column tile, so neither
// grid dimension nor the work-item count of one dispatch (32-bit in the AQL
// packet) can overflow however large the matrix.
#define MHS_ROWS_PER_LAUNCH 32768

// The tiled passes of a launch, as a plan: each item is one kernel instance
// (variant) over a 2-D grid (x: column or source tiles, y: observer rows of
// a chunk of at most MHS_ROWS_PER_LAUNCH) with its parameter block.  Store
// passes write every entry of their rows, add passes add; a launch's add
// items follow all its store items.  launch_all runs the plan as ordinary
// launches; mhs_assemble_batch merges the plans of several launches into one
// persistent launch per variant and stage.
enum Variant {
  V_PAIR_S,      // k_pair<P, 0>: stress family, store
  V_PAIR_D,      // k_pair<P, 1>: displacement, store
  V_FAR64_1,     // k_far (FP64 points), one output, store
  V_FAR64_2,     // k_far (FP64 points), two outputs, store
  V_FARM_1,      // k_far_mixed, one output, store
  V_FARM_2,      // k_far_mixed, two outputs, store
  V_QM_S,        // k_qm<P, true>: stress family, add
  V_QM_D,        // k_qm<P, false>: displacement, add
  V_NEAR,        // k_near<P, 2>: contracted near parts, up to 2 x 2 components, add
  V_NEAR3,       // k_near<P, 3>: the same up to 3 x 3 components, add
  V_COUNT
};
static inline bool is_store(int v) { return v <= V_FARM_2; }

struct PlanItem {
  int var;
  Params pr;
  int64_t gx, gy;
};

template <int P>
static int plan_passes(const mhs_launch* L, const Params& pr, std::vector<PlanItem>& plan) {
  const int64_t n_rows = L->obs_hi - L->obs_lo;
  if (n_rows <= 0 || pr.n_cols <= 0) return 0;
  const int64_t nblk = (pr.n_cols + 255) / 256;
  if (nblk * 256 > 0xffffffffLL) return fail("mhs_assemble: too many columns for one launch");
  const int stages = L->stages ? L->stages : (MHS_STAGE_PAIR | MHS_STAGE_Q | MHS_STAGE_FOLD);
  const int nout = pr.n_a * pr.n_b;
  const bool contracted = use_contracted(pr.kind, pr.sb != nullptr, pr.n_a, pr.n_b, P, pr.flags);
  const bool mixed = pr.flags & F_MIXED;
  const int64_t col_tiles = (pr.n_cols + MHS_ROW_TILE - 1) / MHS_ROW_TILE;
  std::vector<PlanItem> adds;
  for (int64_t lo = L->obs_lo; lo < L->obs_hi; lo += MHS_ROWS_PER_LAUNCH) {
    const int64_t rows = (L->obs_hi - lo) < MHS_ROWS_PER_LAUNCH ? (L->obs_hi - lo) : MHS_ROWS_PER_LAUNCH;
    Params pc = pr;
    pc.obs_lo = lo;
    if (contracted) {
      // output groups: the whole launch when it has one or two outputs, else
      // one receiver frame row a with source rows b0, b0 + 1 per group (the
      // pointers offset to the group, the strides those of the whole launch).
      // The mixed far pass takes one output per group (gw = 1): its two-output
      // instance halves the points per chunk and forms each pair's kappas
      // twice as often, and ran slower than two one-output launches (the
      // three-fault network at order 2: 20.3 ms against 2 x 8.0 ms a chunk).
      const int gw = (mixed && MHS_FARM_SINGLE) ? 1 : 2;
      const bool grouped = nout > gw;
      for (int a = 0; a < (grouped ? pr.n_a : 1); ++a)
        for (int b0 = 0; b0 < (grouped ? pr.n_b : 1); b0 += gw) {
          Params pg = pc;
          if (grouped) {
            pg.rb = pr.rb + a * 3;
            pg.sb = pr.sb + b0 * 3;
            pg.out = pr.out + (a * pr.n_b + b0);
            pg.n_a = 1;
            pg.n_b = pr.n_b - b0 < gw ? pr.n_b - b0 : gw;
          }
          const int gout = pg.n_a * pg.n_b;
          // far pass (writes every entry of the group), then the near pass (adds)
          if (stages & MHS_STAGE_Q) {
            // the mixed mode: a lane per pair at the mixed rule's orders
            const int lpp = mixed ? 1 : (gout == 1 ? MHS_FAR_LPP64 : MHS_FAR_LPP64B);
            const int64_t per = (int64_t)(256 / lpp) * pr.far_tiles;   // sources per block
            const int var = gout == 1 ? (mixed ? V_FARM_1 : V_FAR64_1) : (mixed ? V_FARM_2 : V_FAR64_2);
            plan.push_back({var, pg, (pr.n_src + per - 1) / per, rows});
          }
        }
      // the near pass takes every output of the launch at once (its closed-form
      // tables are formed once per pair and contracted with each output)
      if (stages & MHS_STAGE_PAIR)
        adds.push_back({(pr.n_a <= 2 && pr.n_b <= 2) ? V_NEAR : V_NEAR3, pc, col_tiles, rows});
      continue;
    }
    if (stages & MHS_STAGE_PAIR) {
      if (pr.kind == K_DISP) plan.push_back({V_PAIR_D, pc, col_tiles, rows});
      else plan.push_back({V_PAIR_S, pc, nblk, rows});
    }
    if (pr.kind != K_EIGEN && (stages & MHS_STAGE_Q)) adds.push_back({pr.kind != K_DISP ? V_QM_S : V_QM_D, pc, nblk, rows});
  }
  plan.insert(plan.end(), adds.begin(), adds.end());
  return 0;
}

// One plan item as an ordinary launch.
template <int P>
static int launch_item(const PlanItem& it, hipStream_t st) {
  const dim3 g((unsigned)it.gx, (unsigned)it.gy), b(256);
  constexpr int N64 = MHS_FAR_NPT64;
  constexpr int N64b = N64 > 1 ? N64 / 2 : 1;
  constexpr int L64 = MHS_FAR_LPP64, L64b = MHS_FAR_LPP64B;
  switch (it.var) {
    case V_PAIR_S: hipLaunchKernelGGL((k_pair<P, 0>), g, b, 0, st, it.pr); break;
    case V_PAIR_D: hipLaunchKernelGGL((k_pair<P, 1>), g, b, 0, st, it.pr); break;
    case V_FAR64_1: hipLaunchKernelGGL((k_far<double, P, 1, N64, L64>), g, b, 0, st, it.pr); break;
    case V_FAR64_2: hipLaunchKernelGGL((k_far<double, P, 2, N64b, L64b>), g, b, 0, st, it.pr); break;
    case V_FARM_1: hipLaunchKernelGGL((k_far_mixed<P, 1>), g, b, 0, st, it.pr); break;
    case V_FARM_2: hipLaunchKernelGGL((k_far_mixed<P, 2>), g, b, 0, st, it.pr); break;
    case V_QM_S: hipLaunchKernelGGL((k_qm<P, true>), g, b, 0, st, it.pr); break;
    case V_QM_D: hipLaunchKernelGGL((k_qm<P, false>), g, b, 0, st, it.pr); break;
    case V_NEAR: hipLaunchKernelGGL((k_near<P, 2>), g, dim3(MHS_NEAR_THREADS), 0, st, it.pr); break;
    default: hipLaunchKernelGGL((k_near<P, 3>), g, dim3(MHS_NEAR_THREADS), 0, st, it.pr); break;
  }
  return check(hipGetLastError(), "mhs_assemble: pass launch");
}

// The persistent (queue) kernel of a variant and its resident workgroups per CU.
template <int P>
static const void* queue_kernel(int v) {
  constexpr int N64 = MHS_FAR_NPT64;
  constexpr int N64b = N64 > 1 ? N64 / 2 : 1;
  constexpr int L64 = MHS_FAR_LPP64, L64b = MHS_FAR_LPP64B;
  switch (v) {
    case V_PAIR_S: return (const void*)&kq_pair<P, 0>;
    case V_PAIR_D: return (const void*)&kq_pair<P, 1>;
    case V_FAR64_1: return (const void*)&kq_far<double, P, 1, N64, L64>;
    case V_FAR64_2: return (const void*)&kq_far<double, P, 2, N64b, L64b>;
    case V_FARM_1: return (const void*)&kq_far_mixed<P, 1>;
    case V_FARM_2: return (const void*)&kq_far_mixed<P, 2>;
    case V_QM_S:
    case V_QM_D: return nullptr;          // launched per range (see mhs_assemble_batch)
    case V_NEAR: return (const void*)&kq_near<P, 2>;
    default: return (const void*)&kq_near<P, 3>;
  }
}

template <int P>
static int launch_queue(int v, unsigned grid, const QueueArgs& q, hipStream_t st) {
  const dim3 g(grid), b(256);
  constexpr int N64 = MHS_FAR_NPT64;
  constexpr int N64b = N64 > 1 ? N64 / 2 : 1;
  constexpr int L64 = MHS_FAR_LPP64, L64b = MHS_FAR_LPP64B;
  switch (v) {
    case V_PAIR_S: hipLaunchKernelGGL((kq_pair<P, 0>), g, b, 0, st, q); break;
    case V_PAIR_D: hipLaunchKernelGGL((kq_pair<P, 1>), g, b, 0, st, q); break;
    case V_FAR64_1: hipLaunchKernelGGL((kq_far<double, P, 1, N64, L64>), g, b, 0, st, q); break;
    case V_FAR64_2: hipLaunchKernelGGL((kq_far<double, P, 2, N64b, L64b>), g, b, 0, st, q); break;
    case V_FARM_1: hipLaunchKernelGGL((kq_far_mixed<P, 1>), g, b, 0, st, q); break;
    case V_FARM_2: hipLaunchKernelGGL((kq_far_mixed<P, 2>), g, b, 0, st, q); break;
    case V_QM_S:
    case V_QM_D: return fail("mhs_assemble_batch: the Q-family pass has no queue form");
    case V_NEAR: hipLaunchKernelGGL((kq_near<P, 2>), g, b, 0, st, q); break;
    default: hipLaunchKernelGGL((kq_near<P, 3>), g, b, 0, st, q); break;
  }
  return check(hipGetLastError(), "mhs_assemble_batch: queue launch");
}

// The fold of a launch: pair list, grids and centre loads, near and far points.
template <int P>
static int launch_fold(const mhs_launch* L, const Params& pr, hipStream_t st) {
  const int stages = L->stages ? L->stages : (MHS_STAGE_PAIR | MHS_STAGE_Q | MHS_STAGE_FOLD);
  const int64_t n_rows = L->obs_hi - L->obs_lo;
  if (pr.kind == K_EIGEN || n_rows <= 0 || pr.n_cols <= 0) return 0;
  if (!((stages & MHS_STAGE_FOLD) && (pr.flags & F_FOLD) && L->n_strad > 0)) return 0;
  const bool stress = pr.kind != K_DISP;
  const dim3 block(256);
  if (!L->fold_list || !L->fold_count || !L->fold_err) return fail("mhs_assemble: fold buffers missing");
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, L->fold_count, 2);
  const int64_t nt = n_rows * L->n_strad;
  hipLaunchKernelGGL(k_fold_list, dim3(grid_1d(nt)), dim3(256), 0, st, pr, L->strad,
                     L->n_strad, n_rows, L->fold_list, L->fold_count, L->fold_cap);
  if (check(hipGetLastError(), "k_fold_list launch")) return -1;
  if (!L->fold_work || L->fold_work_items < 1) return fail("mhs_assemble: fold work buffer missing");
  FoldItem* work = (FoldItem*)L->fold_work;
  const int64_t prep_blocks = (L->fold_work_items + 63) / 64;
  hipLaunchKernelGGL((k_fold_prep<P>), dim3((unsigned)(prep_blocks < 4096 ? prep_blocks : 4096)), dim3(64), 0, st,
                     pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items, L->fold_err);
  if (check(hipGetLastError(), "k_fold_prep launch")) return -1;
  const dim3 fg(2048);
  if (stress) {
    hipLaunchKernelGGL((k_fold<P, true, true>), fg, block, 0, st, pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items);
    if (pr.mixed_rule)
      hipLaunchKernelGGL((k_fold<P, true, false, true>), fg, block, 0, st, pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items);
    else
      hipLaunchKernelGGL((k_fold<P, true, false>), fg, block, 0, st, pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items);
  } else {
    hipLaunchKernelGGL((k_fold<P, false, true>), fg, block, 0, st, pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items);
    if (pr.mixed_rule)
      hipLaunchKernelGGL((k_fold<P, false, false, true>), fg, block, 0, st, pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items);
    else
      hipLaunchKernelGGL((k_fold<P, false, false>), fg, block, 0, st, pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items);
  }
  return check(hipGetLastError(), "k_fold launch");
}

template <int P>
static int launch_all(const mhs_launch* L, const Params& pr, hipStream_t st) {
  std::vector<PlanItem> plan;
  if (plan_passes<P>(L, pr, plan)) return -1;
  for (const PlanItem& it : plan)
    if (launch_item<P>(it, st)) return -1;
  return launch_fold<P>(L, pr, st);
}

// Validated parameter block of one launch descriptor.
static int make_params(const mhs_launch* L, Params& pr) {
  if (!L) return fail("mhs_assemble: null launch descriptor");
  if (L->abi_version != MHS_ABI_VERSION) return fail("mhs_assemble: ABI version mismatch");
  if (L->order < 0 || L->order > 2) return fail("mhs_assemble: order must be 0, 1 or 2");
  if (L->kind < K_DISP || L->kind > K_INTERACTION) return fail("mhs_assemble: unknown kind");
  if (!L->obs || !L->src || !L->coef || !L->gl_x || !L->gl_w || !L->gq || !L->out)
    return fail("mhs_assemble: null device pointer");
  if ((L->kind == K_TRACTION || L->kind == K_INTERACTION) && !L->nrm)
    return fail("mhs_assemble: traction/interaction need receiver normals");
  if (L->kind == K_INTERACTION && (!L->rb || L->n_a < 1 || L->n_a > 3 || L->n_b < 1 || L->n_b > 3))
    return fail("mhs_assemble: interaction needs 1..3 receiver and source components");
  if (L->kind == K_INTERACTION && !L->sb && L->n_b != 3)
    return fail("mhs_assemble: identity source basis needs n_b = 3");
  if ((L->flags & F_PARTS) && L->kind != K_STRESS)
    return fail("mhs_assemble: parts needs the stress kind");
  if (L->kind == K_STRESS && (L->flags & (F_PARTS | F_SUBTRACT)) && !L->out2)
    return fail("mhs_assemble: the stress kind with subtract or parts needs out2 for the eigenstress");
  if (L->obs_lo < 0 || L->obs_hi > L->n_obs || L->obs_lo > L->obs_hi)
    return fail("mhs_assemble: bad observer row range");
  const int K = (L->order + 1) * (L->order + 2) / 2;
  if (L->n_cols != L->n_src * K) return fail("mhs_assemble: n_cols must be n_src * K");
  pr.obs = L->obs; pr.nrm = L->nrm; pr.rb = L->rb; pr.sb = L->sb; pr.src = L->src;
  pr.coef = L->coef; pr.gl_x = L->gl_x; pr.gl_w = L->gl_w; pr.gq = L->gq; pr.out = L->out; pr.out2 = L->out2;
  pr.n_obs = L->n_obs; pr.n_src = L->n_src; pr.n_cols = L->n_cols; pr.obs_lo = L->obs_lo;
  pr.out_lo = L->obs_lo;
  pr.near_tiles = L->near_tiles;
  pr.kind = L->kind; pr.flags = L->flags; pr.n_a = L->n_a; pr.n_b = L->n_b;
  pr.mu = L->mu; pr.lam = L->lam; pr.nu = 0.5 * L->lam / (L->lam + L->mu);
  {
    const double c34 = 3.0 - 4.0 * pr.nu;
    pr.C1 = 1.0 / (16.0 * M_PI * pr.mu * (1.0 - pr.nu));
    pr.k5e = -6.0 * (1.0 - pr.nu);
    pr.k7c = 30.0 * (1.0 - pr.nu);
    pr.k7e = 15.0 / (8.0 * M_PI * pr.C1);
    pr.k3 = 1.0 - c34;
    pr.k5c = 3.0 * c34 - 9.0;
  }
  if (L->far_tiles < 0 || L->far_tiles > 4096) return fail("mhs_assemble: far_tiles out of range");
  const bool contracted = use_contracted(L->kind, L->sb != nullptr, L->n_a, L->n_b, L->order, L->flags);
  pr.na_tot = L->n_a;
  pr.nb_tot = L->n_b;
  pr.far_tiles = L->far_tiles ? L->far_tiles : ((contracted && (L->flags & F_MIXED)) ? MHS_FARM_CT : MHS_QC_CT);
  if (fill_rule(L, pr)) return -1;
  return 0;
}

// ----------------------------------------------------------------- batch --
// The merged plan of a batch of launches (mhs_assemble_batch): every plan
// item's parameter block, and per (order, variant) its queue segments in
// launch order.
struct BatchPlan {
  std::vector<Params> params;          // one per plan item
  std::vector<Params> base;            // one per launch (its fold)
  std::vector<QueueSeg> seg[3][V_COUNT];
  int64_t total[3][V_COUNT] = {};
  std::vector<PlanItem> direct[3];     // the Q-family items: ordinary launches
  size_t off_seg[3][V_COUNT] = {};
  size_t off_cnt = 0, bytes = 0;
};

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

static int batch_plan(const mhs_launch* Ls, int n, BatchPlan& bp) {
  if (n < 0 || (n > 0 && !Ls)) return fail("mhs_assemble_batch: bad launch array");
  for (int i = 0; i < n; ++i) {
    const mhs_launch* L = Ls + i;
    Params pr;
    if (make_params(L, pr)) return -1;
    bp.base.push_back(pr);
    std::vector<PlanItem> plan;
    int rc = L->order == 0 ? plan_passes<0>(L, pr, plan) : L->order == 1 ? plan_passes<1>(L, pr, plan)
                                                                          : plan_passes<2>(L, pr, plan);
    if (rc) return -1;
    for (const PlanItem& it : plan) {
      const int64_t tiles = it.gx * it.gy;
      if (tiles <= 0) continue;
      if (it.var == V_QM_S || it.var == V_QM_D) {
        bp.direct[L->order].push_back(it);
        continue;
      }
      if (it.gx > 0x7fffffffLL || it.gy > 0x7fffffffLL) return fail("mhs_assemble_batch: grid too large");
      std::vector<QueueSeg>& sg = bp.seg[L->order][it.var];
      sg.push_back({bp.total[L->order][it.var], (int32_t)bp.params.size(), (int32_t)it.gx});
      bp.total[L->order][it.var] += tiles;
      bp.params.push_back(it.pr);
    }
  }
  if (bp.params.size() > 0x7fffffff) return fail("mhs_assemble_batch: too many plan items");
  size_t off = align256(bp.params.size() * sizeof(Params));
  for (int P = 0; P < 3; ++P)
    for (int v = 0; v < V_COUNT; ++v) {
      bp.off_seg[P][v] = off;
      off = align256(off + bp.seg[P][v].size() * sizeof(QueueSeg));
    }
  bp.off_cnt = off;
  bp.bytes = off + 3 * V_COUNT * sizeof(unsigned long long);
  return 0;
}

// Resident workgroups of each queue kernel per CU, and the CU count, per device.
static std::mutex g_occ_mu;
static int g_occ[MHS_MAX_DEVICES][3][V_COUNT];
static int g_cus[MHS_MAX_DEVICES];

static int queue_grid(int P, int v, int64_t total, unsigned* grid) {
  int dev = 0;
  if (check(hipGetDevice(&dev), "mhs_assemble_batch: hipGetDevice")) return -1;
  if (dev < 0 || dev >= MHS_MAX_DEVICES) return fail("mhs_assemble_batch: device index out of range");
  std::lock_guard<std::mutex> lock(g_occ_mu);
  if (!g_cus[dev]) {
    int cus = 0;
    if (check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute"))
      return -1;
    g_cus[dev] = cus > 0 ? cus : 1;
  }
  if (!g_occ[dev][P][v]) {
    int nb = 0;
    const void* fn = P == 0 ? queue_kernel<0>(v) : P == 1 ? queue_kernel<1>(v) : queue_kernel<2>(v);
    if (check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0), "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
      return -1;
    g_occ[dev][P][v] = nb > 0 ? nb : 1;
  }
  const int64_t resident = (int64_t)g_occ[dev][P][v] * g_cus[dev];
  *grid = (unsigned)(total < resident ? total : resident);
  return 0;
}

// -------------------------------------------------------- static scratch --
// The fold's work buffers for callers that cannot allocate
// (mhs_assemble_static): one file-scope device array per GPU, holding the
// pair counters, the folded-pair list and the folded pairs' grids for up to
// MHS_STATIC_FOLD_PAIRS pairs.  Host code never passes the array's name to
// a kernel (on the host it is a stand-in, not a device address): its device
// address is looked up once per device with hipGetSymbolAddress, with that
// device current, and that address is what the kernels receive.
#define MHS_STR_(x) #x
#define MHS_STR(x) MHS_STR_(x)
static constexpr size_t kStaticListOff = 256;   // after the two int32 pair counters
static constexpr size_t kStaticWorkOff =
    (kStaticListOff + 16 * (size_t)MHS_STATIC_FOLD_PAIRS + 255) / 256 * 256;
static constexpr size_t kStaticBytes = kStaticWorkOff + sizeof(FoldItem) * (size_t)MHS_STATIC_FOLD_PAIRS;
static_assert(kStaticBytes <= (80u << 20), "static scratch stays modest");
static_assert(alignof(FoldItem) <= 256, "fold items align inside the scratch");
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[MHS_MAX_DEVICES];

// The current device's copy of the static scratch (looked up on its first
// use on that device).
static int static_scratch(unsigned char** out) {
  int dev = 0;
  if (check(hipGetDevice(&dev), "mhs_assemble_static: hipGetDevice")) return -1;
  if (dev < 0 || dev >= MHS_MAX_DEVICES) return fail("mhs_assemble_static: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "mhs_assemble_static: hipGetSymbolAddress")) return -1;
    if (!p) return fail("mhs_assemble_static: no device address for the static scratch");
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// ----------------------------------------------------------------- C-ABI --
extern "C" {

const char* mhs_last_error(void) { return g_err; }

int mhs_abi_version(void) { return MHS_ABI_VERSION; }

int mhs_ncoef(void) { return MHS_NCOEF; }

int64_t mhs_fold_item_bytes(int order) {
  (void)order;
  return (int64_t)sizeof(FoldItem);
}

int mhs_src_stride(void) { return MHS_SRC_STRIDE; }

int mhs_fill_material(double mu, double lam, double* host_coef) {
  if (!(mu > 0.0) || !(lam > -2.0 * mu / 3.0)) return fail("mhs_fill_material: inadmissible (mu, lam)");
  if (!host_coef) return fail("mhs_fill_material: null output");
  mhs_fill_coeffs(mu, lam, host_coef);
  return 0;
}

int mhs_assemble(const mhs_launch* L, void* stream) {
  Params pr;
  if (make_params(L, pr)) return -1;
  hipStream_t st = (hipStream_t)stream;
  switch (L->order) {
    case 0: return launch_all<0>(L, pr, st);
    case 1: return launch_all<1>(L, pr, st);
    default: return launch_all<2>(L, pr, st);
  }
}

int64_t mhs_batch_workspace_bytes(const mhs_launch* launches, int n) {
  BatchPlan bp;
  if (batch_plan(launches, n, bp)) return -1;
  return (int64_t)bp.bytes;
}

int mhs_assemble_batch(const mhs_launch* launches, int n, void* dev_ws, void* host_ws, int64_t ws_bytes,
                       void* stream) {
  BatchPlan bp;
  if (batch_plan(launches, n, bp)) return -1;
  if (n == 0) return 0;
  if (!dev_ws || !host_ws) return fail("mhs_assemble_batch: null workspace");
  if (ws_bytes < (int64_t)bp.bytes) return fail("mhs_assemble_batch: workspace smaller than mhs_batch_workspace_bytes");
  if (((uintptr_t)dev_ws | (uintptr_t)host_ws) & 255) return fail("mhs_assemble_batch: workspaces must be 256-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  unsigned char* h = (unsigned char*)host_ws;
  unsigned char* d = (unsigned char*)dev_ws;
  memcpy(h, bp.params.data(), bp.params.size() * sizeof(Params));
  for (int P = 0; P < 3; ++P)
    for (int v = 0; v < V_COUNT; ++v)
      if (!bp.seg[P][v].empty())
        memcpy(h + bp.off_seg[P][v], bp.seg[P][v].data(), bp.seg[P][v].size() * sizeof(QueueSeg));
  memset(h + bp.off_cnt, 0, 3 * V_COUNT * sizeof(unsigned long long));
  // the tables and zeroed counters in one copy (no memset node under capture)
  if (check(hipMemcpyAsync(d, h, bp.bytes, hipMemcpyHostToDevice, st), "mhs_assemble_batch: table upload")) return -1;
  // store passes of every launch, then add passes: one persistent launch per
  // (order, variant) present
  for (int stage = 0; stage < 2; ++stage)
    for (int P = 0; P < 3; ++P)
      for (int v = 0; v < V_COUNT; ++v) {
        if (is_store(v) != (stage == 0) || bp.total[P][v] == 0) continue;
        QueueArgs q;
        q.params = (const Params*)d;
        q.seg = (const QueueSeg*)(d + bp.off_seg[P][v]);
        q.counter = (unsigned long long*)(d + bp.off_cnt) + (P * V_COUNT + v);
        q.total = bp.total[P][v];
        q.n_seg = (int32_t)bp.seg[P][v].size();
        unsigned grid = 0;
        if (queue_grid(P, v, q.total, &grid)) return -1;
        const int rc = P == 0 ? launch_queue<0>(v, grid, q, st) : P == 1 ? launch_queue<1>(v, grid, q, st)
                                                                         : launch_queue<2>(v, grid, q, st);
        if (rc) return -1;
      }
  for (int P = 0; P < 3; ++P)
    for (const PlanItem& it : bp.direct[P]) {
      const int rc = P == 0 ? launch_item<0>(it, st) : P == 1 ? launch_item<1>(it, st) : launch_item<2>(it, st);
      if (rc) return -1;
    }
  // each launch's fold (a device-built pair list of its own)
  for (int i = 0; i < n; ++i) {
    const mhs_launch* L = launches + i;
    const int rc = L->order == 0 ? launch_fold<0>(L, bp.base[i], st) : L->order == 1 ? launch_fold<1>(L, bp.base[i], st)
                                                                                      : launch_fold<2>(L, bp.base[i], st);
    if (rc) return -1;
  }
  return 0;
}

int mhs_assemble_static(const mhs_launch* L, void* stream) {
  if (!L) return fail("mhs_assemble_static: null launch descriptor");
  const bool folds = (L->flags & F_FOLD) && L->n_strad > 0 &&
                     (L->stages == 0 || (L->stages & MHS_STAGE_FOLD)) && L->kind != K_EIGEN;
  mhs_launch S = *L;
  if (folds) {
    if (!L->fold_err) return fail("mhs_assemble_static: the fold needs fold_err (one device int32, an output)");
    if (L->fold_work_items < 0 || L->fold_work_items > MHS_STATIC_FOLD_PAIRS)
      return fail("mhs_assemble_static: the launch's folded (receiver, source) pairs (fold_work_items) exceed "
                  "the static scratch's " MHS_STR(MHS_STATIC_FOLD_PAIRS) "; launch fewer observer rows per call");
    unsigned char* base = nullptr;
    if (static_scratch(&base)) return -1;
    S.fold_count = (int32_t*)base;
    S.fold_list = (int64_t*)(base + kStaticListOff);
    S.fold_work = base + kStaticWorkOff;
    S.fold_cap = MHS_STATIC_FOLD_PAIRS;
    S.fold_work_items = MHS_STATIC_FOLD_PAIRS;
    // the status of this call only (a fill kernel: no memset node under capture)
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, (hipStream_t)stream, L->fold_err, 1);
    if (check(hipGetLastError(), "mhs_assemble_static: clearing fold_err")) return -1;
  }
  return mhs_assemble(&S, stream);
}

int mhs_gather(const double* scratch, double* out, int64_t n_lead, int64_t n_elem_cols, int64_t n_nodes,
               int tail, const int64_t* ptr, const int64_t* idx, const double* sb, int n_a, int n_b,
               void* stream) {
  if (!scratch || !out || !ptr || !idx) return fail("mhs_gather: null device pointer");
  const int64_t total = sb ? n_lead * n_nodes * n_a * n_b : n_lead * n_nodes * tail;
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_gather, dim3(grid_1d(total)), dim3(256), 0, (hipStream_t)stream, scratch, out,
                     n_lead, n_elem_cols, n_nodes, tail, ptr, idx, sb, n_a, n_b);
  return check(hipGetLastError(), "k_gather launch");
}

int mhs_subtract(double* a, const double* b, int64_t n, void* stream) {
  if (!a || !b) return fail("mhs_subtract: null device pointer");
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_sub, dim3(grid_1d(n)), dim3(256), 0, (hipStream_t)stream, a, b, n);
  return check(hipGetLastError(), "k_sub launch");
}

int mhs_strain(double* buf, int64_t n_obs, int64_t n_cols, double mu, double lam, void* stream) {
  if (!buf) return fail("mhs_strain: null device pointer");
  const int64_t total = n_obs * n_cols * 3;
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_strain, dim3(grid_1d(total)), dim3(256), 0, (hipStream_t)stream, buf,
                     n_obs, n_cols, mu, lam);
  return check(hipGetLastError(), "k_strain launch");
}

int mhs_source_major(const double* in, double* out, int64_t n_rec, int64_t n_src, int n_a, int n_b,
                     void* stream) {
  if (!in || !out) return fail("mhs_source_major: null device pointer");
  if (n_a < 1 || n_a > 3 || n_b < 1 || n_b > 3) return fail("mhs_source_major: n_a, n_b must be in 1..3");
  if (n_rec < 0 || n_src < 0) return fail("mhs_source_major: negative size");
  if (n_rec == 0 || n_src == 0) return 0;
  const int nab = n_a * n_b, ts = MHS_TR_LD / nab;
  const int64_t gx = (n_src + ts - 1) / ts, gy = (n_rec + MHS_TR_R - 1) / MHS_TR_R;
  if (gx * 256 > 0xffffffffLL || gy > 0x7fffffffLL) return fail("mhs_source_major: matrix too large for one launch");
  hipLaunchKernelGGL(k_source_major, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, (hipStream_t)stream, in, out,
                     n_rec, n_src, n_a, n_b);
  return check(hipGetLastError(), "k_source_major launch");
}

int mhs_fold_histogram(const mhs_launch* L, unsigned long long* counts, void* stream) {
  if (!L || !counts) return fail("mhs_fold_histogram: null pointer");
  Params pr;
  if (make_params(L, pr)) return -1;
  if (!(L->flags & F_FOLD) || L->n_strad <= 0 || L->kind == K_EIGEN) return 0;
  if (!L->fold_list || !L->fold_count || !L->fold_work || L->fold_work_items < 1)
    return fail("mhs_fold_histogram: fold buffers missing");
  const FoldItem* work = (const FoldItem*)L->fold_work;
  hipStream_t st = (hipStream_t)stream;
  const dim3 g(2048), b(256);
  switch (L->order) {
    case 0: hipLaunchKernelGGL((k_fold_hist<0>), g, b, 0, st, pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items, counts); break;
    case 1: hipLaunchKernelGGL((k_fold_hist<1>), g, b, 0, st, pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items, counts); break;
    default: hipLaunchKernelGGL((k_fold_hist<2>), g, b, 0, st, pr, L->fold_list, L->fold_count, L->fold_cap, work, L->fold_work_items, counts); break;
  }
  return check(hipGetLastError(), "k_fold_hist launch");
}

int mhs_histogram(const mhs_launch* L, unsigned long long* counts, void* stream) {
  if (!L || !counts || !L->obs || !L->src) return fail("mhs_histogram: null pointer");
  Params pr{};
  pr.obs = L->obs; pr.src = L->src; pr.n_obs = L->n_obs; pr.n_src = L->n_src; pr.n_cols = L->n_cols;
  pr.obs_lo = L->obs_lo; pr.flags = L->flags;
  pr.kind = L->kind; pr.sb = L->sb; pr.n_a = L->n_a; pr.n_b = L->n_b;
  if (fill_rule(L, pr)) return -1;
  const int64_t nblk = (L->n_cols + 255) / 256;
  if (L->obs_hi <= L->obs_lo || nblk <= 0) return 0;
  if (nblk * 256 > 0xffffffffLL) return fail("mhs_histogram: too many columns for one launch");
  for (int64_t lo = L->obs_lo; lo < L->obs_hi; lo += MHS_ROWS_PER_LAUNCH) {
    const int64_t rows = (L->obs_hi - lo) < MHS_ROWS_PER_LAUNCH ? (L->obs_hi - lo) : MHS_ROWS_PER_LAUNCH;
    pr.obs_lo = lo;
    hipLaunchKernelGGL(k_hist, dim3((unsigned)nblk, (unsigned)rows), dim3(256), 0, (hipStream_t)stream, pr,
                       L->order, counts);
    if (check(hipGetLastError(), "k_hist launch")) return -1;
  }
  return 0;
}

}  // extern "C"
