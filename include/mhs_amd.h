/* C-ABI of the MI355X mollified Mindlin half-space assembler.
 *
 * The reference (brendanjmeade/mhs) has no FFI: its boundary is the Python
 * public API of src/mhs/matrices.py, re-exported at src/mhs/__init__.py:45-72,
 * whose contract is that the assembly layer below "only ever see[s] contiguous
 * float64 arrays and plain floats" (matrices.py:3-7), and whose designed seam
 * for a compiled replacement is assemble.fullspace_blocks and the assemble_*
 * functions (kernels/assemble.py:18-21, 96-105).  This header IS that seam:
 * plain pointers and sizes, device pointers for every array, no exceptions
 * across the boundary.  Validation, shapes, the byte ceiling and out= stay in
 * the Python layer (mhs_amd/api.py), exactly as in matrices.py.
 *
 * Every entry point returns 0 on success and a negative code on failure;
 * mhs_last_error() then holds a thread-local message.  Launches are
 * asynchronous on the given HIP stream (NULL = the default stream).
 */
#ifndef MHS_AMD_H
#define MHS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHS_ABI_VERSION 6

#define MHS_GQ_TOTAL 2377760   /* sum of n^2 for n = 1..192 */

/* Output kinds: the five public matrix forms (matrices.py:231-580).
 *   MHS_DISP         (n_obs, 3, n_cols, 3)     disp_matrix
 *   MHS_STRESS       (n_obs, 3, 3, n_cols, 3)  total stress in out; with MHS_SUBTRACT or
 *                                              MHS_PARTS the eigenstress in out2 (elastic
 *                                              stress = out - out2 by mhs_subtract, after
 *                                              any gather: stress/total_stress/elastic_strain)
 *   MHS_EIGEN        (n_obs, 3, 3, n_cols, 3)  eigenstress_matrix (the stress form's out2)
 *   MHS_TRACTION     (n_obs, 3, n_cols, 3)     traction_matrix
 *   MHS_INTERACTION  (n_obs, n_cols, n_a, n_b) interaction_matrix */
enum { MHS_DISP = 0, MHS_STRESS = 1, MHS_EIGEN = 2, MHS_TRACTION = 3, MHS_INTERACTION = 4 };

/* Flags: fold surface-reaching sources (fold=True), subtract the eigenstress
 * (subtract_eigenstress=True; the stress form writes it to out2 for the
 * caller's mhs_subtract, the traction and interaction forms subtract it in
 * the kernel), also write the eigenstress to out2 (parts=True),
 * mixed precision (MHS_MIXED: far-field quadrature points in FP32 with FP64
 * accumulation; near field, fold and eigenstress stay FP64). */
enum { MHS_FOLD = 1, MHS_SUBTRACT = 2, MHS_PARTS = 4, MHS_MIXED = 8 };

/* Passes of one assembly: the pair pass writes every entry (direct + image
 * R-family or folded mirror + eigenstress), the Q-family and the Boussinesq
 * fold passes add to it.  A partial mask is for timing a pass alone. */
enum { MHS_STAGE_PAIR = 1, MHS_STAGE_Q = 2, MHS_STAGE_FOLD = 4 };

/* The mixed mode's far-field rule (every form, contracted and tensor passes
 * and the fold's far points).  With MHS_MIXED a (receiver, source) pair is near --
 * closed forms in FP64 -- when sqrt(|x - c|^2 + eps^2) <= near_ratio L (the
 * reference's D* is 10); otherwise each far-field family (direct on the
 * source triangle; image R and Q on the mirrored one, measured from its
 * centroid) takes the Gauss order order[b] + (nodal order) of the first
 * band b with that distance <= band_hi[b] L (the last band takes the rest),
 * squared for the collapsed triangle rule; image Q pairs that are near keep
 * the budget law.  near_ratio = 0 selects the library's default rule: near
 * 4 L, orders 5 / 4 / 3 / 2 up to 6 / 16 / 100 L / beyond, whose quadrature
 * error is below 1e-7 of each pair's own largest entry for every form.  A
 * caller may pass a cheaper rule where its rows tolerate it (the on-fault
 * interaction, whose rows each hold their own element's near field); the
 * mixed mode's bound is 1e-6 of each row's largest entry.  The all-FP64
 * mode always uses the reference's rule (fullspace/defaults.py:20-23,
 * kernels/image.py:249-254).  fold_near_ratio sets the same switch for the
 * fold's surface points alone (closed form within it, the bands' Gauss
 * orders beyond); 0 takes near_ratio. */
#define MHS_RULE_BANDS 8
typedef struct mhs_rule {
  double near_ratio;
  double band_hi[MHS_RULE_BANDS];   /* ascending */
  int32_t order[MHS_RULE_BANDS];    /* 1 .. 190 */
  double fold_near_ratio;           /* 0 or in [1, 10] */
} mhs_rule;

typedef struct mhs_launch {
  int32_t abi_version;      /* MHS_ABI_VERSION */
  int32_t kind;             /* MHS_DISP .. MHS_INTERACTION */
  int32_t order;            /* nodal order 0, 1, 2 (K = 1, 3, 6 columns per source) */
  int32_t flags;            /* MHS_FOLD | MHS_SUBTRACT | MHS_PARTS | MHS_MIXED */
  int32_t n_a, n_b;         /* interaction receiver / source component counts */
  int64_t n_obs, n_src, n_cols;   /* n_cols = n_src * K (element-major columns) */
  int64_t obs_lo, obs_hi;   /* observer rows [obs_lo, obs_hi) of this launch; out row 0 = obs_lo */
  double mu, lam;           /* Material(mu, lam) */
  const double* obs;        /* n_obs x 3 observers (z <= 0) */
  const double* nrm;        /* n_obs x 3 unit receiver normals (traction, interaction) */
  const double* rb;         /* n_obs x n_a x 3 receiver frame rows (interaction) */
  const double* sb;         /* n_src x n_b x 3 source frame rows, or NULL = Cartesian (n_b = 3) */
  const double* src;        /* n_src x mhs_src_stride() per-source records (mhs_amd/geometry.py) */
  const double* coef;       /* mhs_ncoef() image-kernel coefficients from mhs_fill_material */
  const double* gl_x;       /* Gauss-Legendre nodes on [0, 1] for orders 1..192, order n at n(n-1)/2 */
  const double* gl_w;       /* matching weights */
  const double* gq;         /* collapsed Gauss rules on the triangle, order n's n^2 points at
                               (n-1) n (2n-1) / 6, structure of arrays: x1 | x2 | w, each
                               MHS_GQ_TOTAL long (x1 = u_i, x2 = u_j (1 - u_i),
                               w = w_i w_j (1 - u_i) of the order-n Gauss-Legendre rule) */
  double* out;              /* output rows [obs_lo, obs_hi) in the kind's layout */
  double* out2;             /* eigenstress part (MHS_STRESS with MHS_SUBTRACT or MHS_PARTS),
                               same layout, or NULL */
  const int64_t* strad;     /* indices of the straddling (fold) sources, n_strad of them */
  int64_t n_strad;
  int64_t* fold_list;       /* scratch: 2 x fold_cap int64 (receiver, source) pairs */
  int32_t* fold_count;      /* scratch: 2 int32 */
  int32_t* fold_err;        /* 1 int32: bit 0 grid overflow, bit 1 list overflow, bit 2 work
                               buffer overflow (caller zeroes) */
  int64_t fold_cap;
  int32_t stages;           /* 0 = all passes; else a mask of MHS_STAGE_* (timing, serving) */
  int32_t far_tiles;        /* column tiles per block of the contracted far pass (0 = 16);
                               only the grouping of pairs into blocks, never the values */
  void* fold_work;          /* scratch: fold_work_items x mhs_fold_item_bytes(order) bytes, the
                               folded pairs' polar grids and centre loads (prepared in parallel) */
  int64_t fold_work_items;  /* capacity in folded pairs; more pairs than this set fold_err bit 2 */
  mhs_rule rule;            /* MHS_MIXED: the far-field rule (near_ratio 0: the default) */
  const double* near_tiles; /* optional (NULL): per 256 element columns, 8 doubles: the bounding
                               box (lo xyz, hi xyz) of the sources' centroids and mirrored
                               centroids, their largest L, and 1 if any straddles (fold) else
                               0; lets the contracted near pass skip column steps with no near
                               pair (no value changes) */
} mhs_launch;

/* Assemble one matrix block: replaces the per-source loops of
 * src/mhs/kernels/assemble.py:152-331 (assemble_halfspace_disp,
 * assemble_halfspace_total_stress, assemble_eigenstress,
 * assemble_halfspace_traction, assemble_interaction) for the observer rows
 * [obs_lo, obs_hi).  Columns are element-major; the continuous layout's
 * shared columns are summed afterwards by mhs_gather. */
int mhs_assemble(const mhs_launch* launch, void* stream);

/* Several assembly launches at once (the serving loop's batches): replaces
 * the reference's pool of worker processes that each assemble one block of
 * a matrix (src/mhs/parallel.py:132-221, by_source / by_rows), here for
 * blocks of different matrices of uneven size.  Every tiled pass of every
 * launch goes into a device work queue of (launch, observer-row tile,
 * column or source tile) items, one queue per kernel instance, and each
 * queue is drained by ONE persistent launch sized to the resident
 * workgroups (store passes first, then add passes; the uncontracted
 * Q-family pass keeps an ordinary launch per range: compiled with its
 * parameter block in memory its point loop spills); then each launch's
 * fold.
 * The results are bitwise those of mhs_assemble on each launch.  The
 * caller supplies two 256-byte aligned workspaces of at least
 * mhs_batch_workspace_bytes(launches, n): host_ws pinned host memory, which
 * the call fills and copies to dev_ws on the stream (both must stay
 * untouched until the stream has passed the call), dev_ws device memory.
 * No allocation; capturable in a HIP graph (the counters are cleared by the
 * copy, not by a memset). */
int64_t mhs_batch_workspace_bytes(const mhs_launch* launches, int n);
int mhs_assemble_batch(const mhs_launch* launches, int n, void* dev_ws, void* host_ws,
                       int64_t ws_bytes, void* stream);

/* mhs_assemble for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It has no counterpart in the reference.  The fold's work buffers come from
 * a static device array of the library (one per GPU, about 65 MB, held until
 * the process ends) instead of the descriptor: fold_list, fold_count,
 * fold_cap and fold_work are ignored; fold_work_items is the caller's bound
 * on the folded (receiver, source) pairs of the launch's rows and must not
 * exceed MHS_STATIC_FOLD_PAIRS (else the call fails with a message: launch
 * fewer rows per call); fold_err is required when the fold runs and is
 * cleared by the call (its bits then report this call only).  No hipMalloc,
 * hipFree or allocating library call.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other host
 * threads) race and must be serialised by the caller. */
#define MHS_STATIC_FOLD_PAIRS 3200
int mhs_assemble_static(const mhs_launch* launch, void* stream);

/* Continuous layout scatter-add (assemble.py:69-93, 328-330), as a gather:
 * out[..., g, ...] = sum over the element columns idx[ptr[g]:ptr[g+1]]
 * (ascending) of scratch[..., c, ...].  With sb (n_nodes x n_b x 3) the
 * scratch is an interaction block with Cartesian slip (n_lead, n_elem_cols,
 * n_a, 3) and the slip is resolved on node g's own frame rows. */
int mhs_gather(const double* scratch, double* out, int64_t n_lead, int64_t n_elem_cols,
               int64_t n_nodes, int tail, const int64_t* ptr, const int64_t* idx,
               const double* sb, int n_a, int n_b, void* stream);

/* Branch histogram of the launch's pairs (adds into 24 device uint64
 * counters): the FP64 rule's pairs; direct near / far-12^2 / far-8^2; image
 * near / far12 / far8; folded pairs; Q-family points and pairs; folded mirror
 * near / far12 / far8; then, for MHS_MIXED on the contracted interaction
 * (else zero), the mixed rule per (row, source) pair: pairs; direct near /
 * far pairs / far points; image R near / far pairs / far points; Q pairs /
 * points / chunks of the mixed far pass; R chunks.  The work model of the
 * roofline figures. */
int mhs_histogram(const mhs_launch* launch, unsigned long long* counts, void* stream);

/* Point counts of the fold's passes (adds into 4 device uint64 counters),
 * read from the pair list and polar grids that an mhs_assemble of the same
 * descriptor left in its fold buffers: closed-form (near) points, Gauss
 * (far) points and their Gauss points (the order squared), each per node
 * column, and the folded (pair, node) items.  No counterpart in the
 * reference: the work model of the fold's roofline (fold.py:84-380 is the
 * algorithm counted). */
int mhs_fold_histogram(const mhs_launch* launch, unsigned long long* counts, void* stream);

/* a[i] -= b[i] for i < n, device FP64: the elastic stress as the finished
 * total stress minus the finished eigenstress, as matrices.py:287-294 forms
 * it (sig -= star), so stress_matrix == total_stress_matrix -
 * eigenstress_matrix bitwise. */
int mhs_subtract(double* a, const double* b, int64_t n, void* stream);

/* Elastic strain from elastic stress in place (matrices.py:572-579). */
int mhs_strain(double* buf, int64_t n_obs, int64_t n_cols, double mu, double lam, void* stream);

/* Source-major copy of an interaction matrix for consumers that read one
 * source's column at every receiver (studies/active_fractal/kernel.py:14-23,
 * 259-261, to_source_major): out[s][b][a][r] = in[r][s][a][b], in of shape
 * (n_rec, n_src, n_a, n_b), out (n_src, n_b, n_a, n_rec), both contiguous
 * device FP64; 1 <= n_a, n_b <= 3. */
int mhs_source_major(const double* in, double* out, int64_t n_rec, int64_t n_src, int n_a, int n_b,
                     void* stream);

/* Per-triangle geometry of the shared-node table, as shared.py:175-273 and
 * tdcs.py:49-66 compute it: area_normals (n_tri, 3) = area * upward unit
 * normal, area3 (n_tri) = area / 3 (device), and into host stats[4]: the
 * shortest edge, the largest |coordinate|, the number of triangles with a
 * non-finite coordinate and the number of collinear ones.  stats_dev is 32
 * bytes of device memory for the reductions.  Synchronises the stream. */
int mhs_tri_geometry(const double* tris, int64_t n_tri, double* area_normals, double* area3, void* stats_dev,
                     double* stats, void* stream);

/* Workspace bytes of mhs_shared_nodes for n_tri triangles (-1 if out of
 * range: 1 <= 3 n_tri < 2^31). */
int64_t mhs_weld_workspace_bytes(int64_t n_tri);

/* Shared-node table of the continuous layout (shared.py:175-273,
 * shared_nodes) on the device, for large fault networks.  Inputs: tris
 * (n_tri, 3, 3), groups (n_tri) int64, tol = tol_over_h * shortest edge,
 * area_normals (n_tri, 3) = area * upward unit normal and area3 (n_tri) =
 * area / 3, all computed by the caller as the host table computes them.
 * Outputs at capacity, n_v + n_e <= 6 n_tri nodes: ids (n_tri, 6) (v1 v2 v3
 * m12 m23 m31 node ids), coords / normals (nodes, 3), weights (nodes),
 * group (nodes) int64, boundary (nodes) uint8, edges (3 n_tri, 2) welded
 * vertex pairs, counts and n_forward (3 n_tri) int32 (triangles on the edge,
 * forward traversals), n_nodes[2] = {n_v, n_e} and refusals[0] (the
 * MHS_REFUSE_* bits found) in host memory.  Node ids and the order of every
 * sum equal the host table's.  The refusal messages are the caller's, from
 * ids, edges, counts and n_forward, read only when a bit calls for them.
 * Synchronises the stream four times (key ranges, node counts, refusals). */
#define MHS_REFUSE_DEGENERATE 1     /* two corners of a triangle welded */
#define MHS_REFUSE_NON_MANIFOLD 2   /* an edge on more than two triangles */
#define MHS_REFUSE_SHARED 4         /* some edge on exactly two triangles */
#define MHS_REFUSE_WINDING 8        /* a shared edge traversed the same way twice */
int mhs_shared_nodes(const double* tris, const int64_t* groups, int64_t n_tri, double tol,
                     const double* area_normals, const double* area3, void* work, int64_t work_bytes,
                     int64_t* ids, double* coords, double* normals, double* weights, int64_t* group,
                     uint8_t* boundary, int64_t* edges, int* counts, int* n_forward, int64_t* n_nodes,
                     int* refusals, void* stream);

/* Per-material image-kernel coefficients into host memory (mhs_ncoef()
 * doubles): the exact rational image table with nu = lam / (2 (lam + mu)) and
 * the scale 2 (lam + mu) / (pi mu (lam + 2 mu)) folded in
 * (src/mhs/kernels/image.py:95-104, 276-297, 453-456). */
int mhs_fill_material(double mu, double lam, double* host_coef);

/* Bytes of one folded pair's slot in fold_work at this order. */
int64_t mhs_fold_item_bytes(int order);

int mhs_ncoef(void);
int mhs_src_stride(void);
int mhs_abi_version(void);
const char* mhs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
