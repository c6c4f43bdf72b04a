/*
 * libsmx_hip: C ABI of the MI355X (gfx950) demag + LLG hot path.
 *
 * Conventions
 *   - Every pointer argument is a device pointer unless stated otherwise.
 *   - Vector fields are interleaved per node: [x0, y0, z0, x1, ...].
 *   - "SELL-64" sparse storage: rows are grouped in slices of 64 (one
 *     wavefront).  Slice s holds width_s = (slice_ptr[s+1]-slice_ptr[s])/64
 *     entries per row, stored column-major inside the slice: entry j of row
 *     64*s+l sits at position slice_ptr[s] + 64*j + l of `cols` and, with k
 *     value components per entry, at k*(that position) of the value array.
 *     Padding entries carry value 0 and a valid column index.
 *   - Functions are asynchronous on `stream` (a hipStream_t passed as void*),
 *     allocate no device memory (workspaces are caller-provided) and return a
 *     smx_status.  smx_last_error() describes the last failure of the thread.
 *     One exception: the waiting form of the CG solvers (smx_pcg_jacobi,
 *     smx_pcg_amg and their _mixed forms with check_every >= 0) blocks the
 *     calling host thread on
 *     hipStreamSynchronize(stream) between batches of iterations, because it
 *     returns the iteration count and residual in host memory; their
 *     speculative form (check_every < 0) returns without waiting.  The
 *     tiled ACA for large blocks (smx_aca_large) waits likewise every 8 rank
 *     steps.
 *   - Results are deterministic (bitwise, run to run): reductions sum in a
 *     fixed order, and no floating-point atomics are used.  The one atomic
 *     accumulation, the Lindholm row assembly, adds 64-bit fixed-point
 *     integers (units of 2^-60), which is independent of the order.
 *
 * Reference interfaces replaced (paths relative to src/spinfemx of the
 * reference) are named at each entry point.
 */
#ifndef SMX_HIP_H
#define SMX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SMX_OK = 0,
  SMX_ERR_SHAPE = 1,          /* ValueError in the reference               */
  SMX_ERR_VALUE = 2,          /* ValueError                                */
  SMX_ERR_NOT_CONVERGED = 3,  /* KSP did not converge (never checked there) */
  SMX_ERR_OOM = 4,            /* workspace too small / MemoryError          */
  SMX_ERR_HIP = 5,            /* RuntimeError from the device runtime       */
  SMX_ERR_TYPE = 6            /* TypeError (unsupported dtype)              */
} smx_status;

typedef enum { SMX_F64 = 0, SMX_F32 = 1, SMX_BF16 = 2 } smx_dtype;

/* Number of doubles per triangle in the geometry table (structure of arrays:
 * field f of triangle t at geom[f * ntri_total + t]).  Fields:
 *   0-8   p0, p1, p2 (xyz each)       9-11  unit normal n
 *   12-14 s_edge (|p1-p0|, |p2-p1|, |p0-p2|) + 1e-30
 *   15-23 eta_i = n x xi_i (row i)    24-32 gamma_ij = xi_{i+1}.xi_j
 *   33    c = 1 / (8 pi (area + 1e-300))                                   */
#define SMX_TRI_FIELDS 34
#define SMX_LINDHOLM_TILE_COLS 1024

int smx_version(void);
const char* smx_last_error(void);

/* ---------------------------------------------------------------------- *
 * (a) Lindholm double-layer block / row assembly.
 * Replaces LindholmEntryProviderGPU.fill_block_device
 * (gpu/fields/lindholm_hmatrix_gpu/entry_provider_gpu.py:687-750, kernels
 * :205-347) and the numba assemble_B_rows_owned (gpu/fields/Demag_Lindholm.py:51-116).
 *
 * out[r*ld_row + c*ld_col] = sum over relevant triangles i (global index
 * rel_tri[i]) not incident to target rows[r] (when incident_policy == 0) of the
 * Lindholm weight of vertex k of that triangle, for c = rel_cols[3*i+k] >= 0.
 * ncols <= SMX_LINDHOLM_TILE_COLS.  Entries without a contribution are 0.
 * One wavefront per target row; triangle geometry is staged in LDS.
 * ---------------------------------------------------------------------- */
int smx_lindholm_fill(const double* Xb, const int32_t* rows, int64_t nrows,
                      const double* tri_geom, int64_t ntri_total,
                      const int32_t* tri_lid, const int32_t* rel_tri,
                      const int32_t* rel_cols, int64_t nrel, int64_t ncols,
                      double* out, int64_t ld_row, int64_t ld_col,
                      int32_t incident_policy, void* stream);

/* Batched fill: one task per (block, column tile).  Task t fills rows
 * rows[rows_off .. rows_off+nrows) against the relevant triangles
 * rel_tri[rel_off .. rel_off+nrel) (local columns rel_cols[3*(rel_off+i)+k]),
 * writing out[out_off + r*ld_row + c*ld_col].  Workgroup g handles rows
 * wg_row0[g] .. +3 of task wg_task[g].  Replaces the per-block launches of
 * HMatrixBuilderGPU._build_block (lindholm_hmatrix_gpu/builder.py:683-781). */
typedef struct {
  int64_t out_off;
  int64_t ld_row;
  int64_t ld_col;
  int64_t rows_off;
  int64_t rel_off;
  int32_t nrows;
  int32_t nrel;
  int32_t ncols;
  int32_t pad;
} smx_fill_task;

int smx_lindholm_fill_batched(const double* Xb, const double* tri_geom, int64_t ntri_total,
                              const int32_t* tri_lid, const int32_t* rows,
                              const int32_t* rel_tri, const int32_t* rel_cols,
                              const smx_fill_task* tasks, const int32_t* wg_task,
                              const int32_t* wg_row0, int64_t nwg, double* out,
                              int32_t incident_policy, void* stream);

/* Write out[i*ld + i] = diag[i] (the solid-angle jump omega/4pi - 1) and, when
 * out_lo != NULL, convert the n x n FP64 matrix into dtype `lo_dtype` (FP32 or
 * BF16, round-to-nearest-even) with leading dimension ld.  Used after
 * smx_lindholm_fill to finish the dense B (Demag_Lindholm.py:73-75). */
int smx_dense_finish(double* B, int64_t n, int64_t ld, const double* diag,
                     void* out_lo, int32_t lo_dtype, void* stream);

/* ---------------------------------------------------------------------- *
 * (b) Dense boundary matvec y = B x, B row-major n x n (leading dimension ldb)
 * stored in FP64, FP32 or BF16; x, y FP64; FP64 accumulation.  When
 * `gather` != NULL the input is x[k] = src[gather[k]] - (*shift) * shift_scale
 * (the boundary gather xb = u1[bdofs] with the nullspace projection folded in).
 * Replaces cp.matmul(B_dense_gpu, xb) (gpu/fields/Demag_Lindholm_GPU.py:596-616).
 * ---------------------------------------------------------------------- */
int smx_dense_bmatvec(const void* B, int32_t dtype, int64_t n, int64_t ldb,
                      const double* src, const int32_t* gather,
                      const double* shift, double shift_scale, double* xbuf,
                      double* y, void* stream);

/* Multi-RHS dense product Y = B X on the matrix cores (nrhs <= 64 columns,
 * column j of X / Y at X + j*ldx / Y + j*ldy, FP64 in and out).  B as in
 * smx_dense_bmatvec with ldb a multiple of 8 and zero padding.  BF16 B uses
 * v_mfma_f32_16x16x32_bf16 with X rounded to BF16 and FP32 accumulation;
 * FP32 B uses v_mfma_f32_16x16x4_f32 (X rounded to FP32); FP64 B uses
 * v_mfma_f64_16x16x4_f64.  Split-K partials are added in a fixed order in
 * FP64.  work: smx_dense_bmatmul_ws_size(n, dtype) bytes.
 * Batched form of the boundary apply (gpu/fields/Demag_Lindholm_GPU.py:596-616)
 * for many right-hand sides at once (the c2 precision study, FD directions).
 */
int64_t smx_dense_bmatmul_ws_size(int64_t n, int32_t dtype);
int smx_dense_bmatmul(const void* B, int32_t dtype, int64_t n, int64_t ldb,
                      const double* X, int64_t ldx, int32_t nrhs, double* Y,
                      int64_t ldy, void* work, int64_t work_bytes, void* stream);

/* Transposed dense matvec y = B^T x (adjoint of the boundary operator).
 * work must hold (ceil(n/256)) * n doubles. */
int smx_dense_bmatvec_t(const void* B, int32_t dtype, int64_t n, int64_t ldb,
                        const double* x, double* y, double* work, void* stream);

/* ---------------------------------------------------------------------- *
 * (a/b) Full-pivot ACA of many dense blocks, one workgroup per block, no host
 * synchronization.  Replaces fullaca_dense_gpu + rank1_residual_update
 * (lindholm_hmatrix_gpu/compressors.py:251-372) with the algorithm of
 * fullaca_dense (compressors.py:35-99): pivot = first argmax |R| in row-major
 * order; stop if |pivot| <= eps_mach*||A||_F; u = R[:,j], v = R[i,:]/pivot;
 * R -= u v^T (unfused multiply and subtract, as numpy does); stop when
 * ||R||_F/||A||_F <= rtol.  The block (row-major, leading dimension n) is
 * overwritten by the residual.  U is column-major (m x rcap), V row-major
 * (rcap x n).  status: 0 converged, 1 pivot break, 2 reached rcap, 3 zero block.
 * ---------------------------------------------------------------------- */
typedef struct {
  int64_t a_off;
  int64_t u_off;
  int64_t v_off;
  int32_t m;
  int32_t n;
  int32_t rcap;
  int32_t pad;
} smx_aca_task;

int smx_aca_batched(double* A, const smx_aca_task* tasks, int64_t ntasks, double rtol,
                    double* U, double* V, int32_t* rank_out, double* resid_out,
                    int32_t* status_out, void* stream);

/* The same ACA for large blocks, each rank step spread over row tiles of
 * SMX_ACA_TILE_ROWS rows: one launch over the tiles of every open block
 * (rank-1 update, partial ||R||_F^2 and first argmax |R|), then one pivot
 * workgroup per block (the partials in tile order, the stopping tests of
 * smx_aca_batched in its order, the next u and v).  Same per-element
 * arithmetic and pivots as smx_aca_batched; ||R||_F is summed per tile.
 * Tile t covers rows tile_row0[t] .. +SMX_ACA_TILE_ROWS of block
 * tile_task[t]; block b owns tiles tile_first[b] .. tile_first[b+1]-1
 * (consecutive, rows ascending).  max_rcap >= every task's rcap.  Waits on
 * `stream` every 8 rank steps to read how many blocks are done (the waiting
 * form, as the CG solvers).  Replaces the per-step rank1_residual_update
 * loop of fullaca_dense_gpu (lindholm_hmatrix_gpu/compressors.py:251-372)
 * for blocks too large for one workgroup.  work:
 * smx_aca_large_ws_doubles(ntasks, ntiles) doubles. */
#define SMX_ACA_TILE_ROWS 64
int64_t smx_aca_large_ws_doubles(int64_t ntasks, int64_t ntiles);
int smx_aca_large(double* A, const smx_aca_task* tasks, int64_t ntasks, const int32_t* tile_task,
                  const int32_t* tile_row0, const int32_t* tile_first, int64_t ntiles,
                  int32_t max_rcap, double rtol, double* U, double* V, int32_t* rank_out,
                  double* resid_out, int32_t* status_out, double* work, int64_t work_doubles,
                  void* stream);

/* ---------------------------------------------------------------------- *
 * (a/b) Truncated-SVD and randomized-SVD compression of many dense blocks,
 * one workgroup per block, no library calls.  Block, U and V layouts as
 * smx_aca_batched (A row-major m x n at a_off; column r of U, scaled by s_r,
 * at u_off + r m; row r of V at v_off + r n).  Rank rule of the reference
 * (compressors.py:402-406): with s descending, rel_r = sqrt(max(||A||^2 -
 * sum_{i<=r} s_i^2, 0)) / ||A||; rank = first r with rel_r <= rtol, else
 * min(mr, min(m, n)).  status: 0 factors stored, 2 rank above rcap (nothing
 * stored), 3 zero block or mr = 0 (rank 0, rel 0), 4 not resolved (rel inf),
 * 7 (randomized only) the sketch found no rank meeting rtol and s < mr:
 * call again with a wider sketch.  `work` holds smx_svd_ws_doubles(m, n, s)
 * doubles per task at w_off.
 * ---------------------------------------------------------------------- */
#define SMX_SVD_KMAX 128   /* rows of a Jacobi problem: QR steps / sketch width */

typedef struct {
  int64_t a_off;
  int64_t u_off;
  int64_t v_off;
  int64_t w_off;   /* doubles into `work` */
  int64_t om_off;  /* randomized: the n x s test matrix (row-major) in `omega` */
  int32_t m;
  int32_t n;
  int32_t rcap;    /* rank capacity of the U/V slots */
  int32_t mr;      /* usable rank: min(m, n), or min(max_rank, m, n) */
  int32_t s;       /* randomized: sketch width of this call (1..SMX_SVD_KMAX) */
  int32_t pad;
} smx_svd_task;

/* Workspace doubles of one task: s = 0 for smx_svd_batched, the sketch width
 * for smx_rsvd_round; -1 for invalid sizes. */
int64_t smx_svd_ws_doubles(int32_t m, int32_t n, int32_t s);

/* Truncated SVD (replaces truncated_svd_gpu, compressors.py:373-414):
 * Householder QR with column pivoting of A^T, stopped once the trailing
 * Frobenius norm is <= 1e-4 rtol ||A||_F (or after SMX_SVD_KMAX steps, if it
 * is then <= 1e-2 rtol ||A||_F), then a one-sided Jacobi SVD of the leading
 * rows of the triangular factor: captured energies within that floor squared
 * of A's, and the reported residual is that of the stored factors.  A block
 * that does not reach the 1e-2 floor within SMX_SVD_KMAX steps gets status 4.
 * A is overwritten (rows permuted, reflectors stored). */
int smx_svd_batched(double* A, const smx_svd_task* tasks, int64_t ntasks, double rtol, double* U,
                    double* V, double* work, int32_t* rank_out, double* resid_out,
                    int32_t* status_out, void* stream);

/* One round of the adaptive randomized SVD (replaces the loop body of
 * randomized_svd_gpu, compressors.py:417-492): Y = A Omega; power_iterations
 * times Q = orth(Y), Y = A A^T Q; Q = orth(Y); B = Q^T A; SVD of B by
 * one-sided Jacobi; rank rule over the s values.  A is not modified. */
int smx_rsvd_round(const double* A, const smx_svd_task* tasks, int64_t ntasks, const double* omega,
                   double rtol, int32_t power_iterations, double* U, double* V, double* work,
                   int32_t* rank_out, double* resid_out, int32_t* status_out, void* stream);

/* ---------------------------------------------------------------------- *
 * (b) H-matrix boundary matvec y = B x (replaces HMatrixGPUPackedFused.matvec,
 * lindholm_hmatrix_gpu/backend_fused.py:496-512, with the semantics of
 * HMatrixCPUData.matvec_cpu, storage.py:284-299).  Three launches:
 *   xp = src[gidx] - shift                 (perm gather, optional boundary gather)
 *   z[t.z_off + g] = V_t[g, :] . xp[t.c0 + (0..n)]   for g < t.kc
 *                                          (one wavefront per z task t)
 *   yp[p] = d_p xp[p] + Near[p,:] xp + U_L[p - l0, :] . zl_L;  y[perm[p]] = yp[p]
 *                                          (one workgroup per leaf L)
 * With ynear set, Near[p,:] xp is computed by a launch of its own before the
 * z launch, one wavefront per row, and the leaf launch streams U only: this
 * spreads the near field over the whole device when the leaves are few.
 * Layout (every factor load is 16 bytes and aligned):
 *  - a z task covers kc <= SMX_HMAT_Z_GROUP rank rows of one far block's V:
 *    row g at V + v_off + g*ld, ld = n rounded up to SMX_HMAT_V_ALIGN_F32
 *    entries for FP32 factors (rows on 128-byte cache lines) or
 *    SMX_HMAT_V_ALIGN_F64 for FP64 (zero padded), v_off a multiple of it;
 *  - a leaf L (rows leaf_start..leaf_stop, at most 64) owns a dense row-major
 *    matrix of its far rows: row i at U + leaf_u_off[L] + i*leaf_kw[L],
 *    leaf_kw a multiple of 4, zero padded.  Its column k multiplies
 *    zl_L[k] = z[leaf_zg[leaf_zg_off[L] + k]] for k < K_L =
 *    leaf_zg_off[L+1] - leaf_zg_off[L] (0 beyond), staged in LDS
 *    (kw_max <= SMX_HMAT_KW_MAX doubles per workgroup);
 *  - the near rows are CSR with an even entry count per row (pairs padded
 *    with value 0 and a valid column), near_ptr[p] even.
 * No atomics: every z entry and every output row has one producer and a
 * fixed summation order.  Factors FP32 or FP64, near values FP64 or FP32,
 * FP64 accumulation.
 * ---------------------------------------------------------------------- */
#define SMX_HMAT_Z_GROUP 8
#define SMX_HMAT_V_ALIGN_F32 32
#define SMX_HMAT_V_ALIGN_F64 4
#define SMX_HMAT_KW_MAX 16384
#define SMX_HMAT_NEAR_W_MAX 8192

/* A near task: rows p0..p1-1 (permuted) and the w columns they touch, listed
 * at near_stage + s_off. */
typedef struct smx_hmat_ntask_s {
  int64_t s_off;
  int32_t p0;
  int32_t p1;
  int32_t w;
  int32_t pad;
} smx_hmat_ntask;

typedef struct {
  int64_t v_off;
  int64_t z_off;
  int32_t c0;
  int32_t n;
  int32_t kc;
  int32_t pad;
} smx_hmat_ztask;

typedef struct {
  int64_t nb;
  const int32_t* gidx;
  const int32_t* perm;
  const double* diag_p;
  const int64_t* near_ptr;
  const int32_t* near_idx;
  const void* near_val;
  int32_t near_dtype;
  int32_t factor_dtype;
  int64_t nztask;
  const smx_hmat_ztask* ztask;
  const void* V;
  int64_t nleaf;
  const int32_t* leaf_start;
  const int32_t* leaf_stop;
  const int64_t* leaf_u_off;
  const int32_t* leaf_kw;
  const int64_t* leaf_zg_off; /* nleaf + 1 */
  const int32_t* leaf_zg;
  int32_t kw_max;
  int32_t leaf_waves; /* wavefronts per leaf workgroup: 4, 8 or 16 (0: 8 for
                       * FP32 factors, 16 for FP64) */
  const void* U;
  double* xp; /* workspace nb */
  double* z;  /* workspace: sum of ranks */
  /* optional workspace nb (NULL: off): the near rows are summed by a launch
   * of their own (one wavefront per row) into ynear[p], and the leaf launch
   * reads that sum instead of the near row */
  double* ynear;
  /* optional, with ynear (NULL: one wavefront per row over near_idx): the
   * near launch as one workgroup per near task (smx_hmat_ntask).  Task t
   * stages xp[near_stage[s_off + i]] for i < w in LDS and sums its rows
   * p0..p1-1 with near_lidx, the slot of each entry's column in that list
   * (uint16, the layout of near_idx, padding entries a valid slot).  Every
   * row is in one task; tasks with p0 == p1 are empty.  w <= near_w_max <=
   * SMX_HMAT_NEAR_W_MAX.  The same sums in the same order as the
   * wavefront-per-row launch, so the same bits. */
  const struct smx_hmat_ntask_s* near_task;
  int64_t n_near_task;
  const int32_t* near_stage;
  const uint16_t* near_lidx;
  int32_t near_w_max;
  int32_t near_waves; /* wavefronts per near task workgroup: 4, 8 or 16 (0: 8) */
  /* optional (NULL: the leaf launch gathers through leaf_zg): leaf L's list
   * as runs of consecutive z entries, runs leaf_zr_off[L] .. leaf_zr_off[L+1]-1
   * with run q = (leaf_zr[2q], leaf_zr[2q+1]) = (first z entry, first column
   * k), running to the next run's first column (K_L for the last): zl_L[k] =
   * z[leaf_zr[2q] + k - leaf_zr[2q+1]], the same values as leaf_zg gives. */
  const int64_t* leaf_zr_off; /* nleaf + 1 */
  const int32_t* leaf_zr;
} smx_hmatrix;

/* The near rows as dense tiles for smx_hmat_matmat: tile t covers rows
 * row[16 t + s] (permuted indices; s < nrows <= 16, -1 beyond) and the
 * columns col[col_off .. col_off + w) (permuted indices, the union of its
 * rows' near columns); its values sit row-major at val + a_off (val_dtype
 * SMX_F64 or SMX_F32), rows padded to a multiple of 4 entries with zeros,
 * a_off a multiple of 4.  Every row 0..nb-1 is in exactly one tile.  With
 * tiles the rows are finished in permuted order and written to Y by an LDS
 * transpose through iperm (contiguous stores), not scattered per entry. */
typedef struct {
  int64_t a_off;
  int64_t col_off;
  int32_t w;
  int32_t nrows;
} smx_hmat_ntile;

typedef struct {
  int64_t ntile;
  const smx_hmat_ntile* tile;
  const int32_t* row; /* 16 * ntile */
  const int32_t* col;
  const void* val;
  const int32_t* iperm; /* inverse of H->perm: the permuted row of output index i */
  int32_t val_dtype;
  int32_t pad;
} smx_hmat_near_tiles;

/* Y = B X for nrhs <= 64 right-hand sides (column j of X at X + j*ldx, of
 * Y at Y + j*ldy, indexed through H->gidx like smx_hmat_matvec's src, no
 * shift), with the low-rank products on the matrix cores: the z stage
 * Z = V_b X_b and the leaf stage U_L Z_L become 16 x 16 MFMA tiles over the
 * 64 columns (FP32 factors: v_mfma_f32_16x16x4_f32 with X and Z rounded to
 * FP32 once; FP64: v_mfma_f64_16x16x4_f64); the near field and the diagonal
 * sum in FP64 over the staged X: with NT (optional) as v_mfma_f64_16x16x4_f64
 * tiles over NT's dense near tiles, else one wavefront per CSR row.  No counterpart call in the reference, whose matvec takes one
 * vector (backend_fused.py:496-512): it replaces a loop of matvecs.  nz: the
 * packed operator's z entries (the largest z_off + kc); work: at least
 * smx_hmat_matmat_ws_size(H, nz) bytes.  H's z tasks here hold up to
 * SMX_HMAT_MM_ROWS rank rows (kc; one MFMA row tile: rows beyond it are not
 * computed).  With FP32 factors a leaf's four wavefronts split its columns
 * in four ranges, each covering all of the leaf's rows, and their partial
 * tiles are added in FP64 in wavefront order. */
#define SMX_HMAT_MM_ROWS 16
int64_t smx_hmat_matmat_ws_size(const smx_hmatrix* H, int64_t nz);
int smx_hmat_matmat(const smx_hmatrix* H, const smx_hmat_near_tiles* NT, int64_t nz, const double* X,
                    int64_t ldx, int32_t nrhs, double* Y, int64_t ldy, void* work, int64_t work_bytes,
                    void* stream);
int smx_hmat_matvec(const smx_hmatrix* H, const double* src, const double* shift,
                    double shift_scale, double* y, void* stream);

/* The H-matrix matvec without caller workspaces.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same operator, arguments and result as smx_hmat_matvec (the
 * same kernels), except that H->xp, H->z and H->ynear are not used: the
 * gathered x, the z entries and the near sums live in a fixed array in
 * device memory that the library holds on each GPU it runs on until the
 * process ends (H->ynear != NULL still selects the near rows in a launch of
 * their own).  nz is the number of z entries (the sum of the far blocks'
 * ranks).  Needs nb + nz (+ nb with ynear) <= SMX_HMAT_NOALLOC_MAX_DOUBLES
 * (8 MiB of doubles; the 83 202-node film needs about 0.62 M), else
 * SMX_ERR_SHAPE before any launch.  Calls nothing that allocates.  Every
 * call on one GPU uses the same array, so calls that overlap on one GPU
 * (other streams or host threads) race; every scratch entry a call reads is
 * written earlier in the same call, so nothing needs clearing between calls. */
#define SMX_HMAT_NOALLOC_MAX_DOUBLES (1 << 20)
int smx_hmat_matvec_noalloc(const smx_hmatrix* H, int64_t nz, const double* src,
                            const double* shift, double shift_scale, double* y, void* stream);

/* ---------------------------------------------------------------------- *
 * (c) SELL-64 sparse kernels (replace PETSc AIJCUSPARSE MatMult on Div, A_u2_raw,
 * G and K_lin: Demag_Lindholm_GPU.py:570, 632, 662; llg_module_GPU.py:963-964).
 * kind 0: y_i = sum_j a_ij x_j                     (k = 1)
 * kind 1: y_i = sum_j a_ij . x_j   (x 3-vector)    (k = 3, divergence)
 * kind 2: y_i = sum_j a_ij x_j     (y 3-vector)    (k = 3, gradient)
 * kind 3: y_i = sum_j A_ij x_j     (3x3 blocks)    (k = 9)
 * When partials != NULL (kind 0/1), block b of the launch writes the sum of
 * its rows of y to partials[b]; smx_sell_spmv_blocks(nrows) blocks are used.
 * ---------------------------------------------------------------------- */
int smx_sell_spmv(int64_t nrows, const int64_t* slice_ptr, const int32_t* cols,
                  const double* vals, int32_t kind, const double* x, double* y,
                  double* partials, void* stream);
int64_t smx_sell_spmv_blocks(int64_t nrows);

/* Doubles needed by a reduction workspace for n rows (partials + counter). */
int64_t smx_reduce_ws_size(int64_t n);

/* ---------------------------------------------------------------------- *
 * (c) Jacobi-preconditioned CG (Chronopoulos-Gear form: one fused
 * SpMV + update + 4-dot kernel per iteration, device-side convergence test,
 * fixed-order reductions without atomics).  Replaces PETSc KSP cg (+gamg) of
 * the two Poisson solves (Demag_Lindholm_GPU.py:216-226, 271-281, 569-640).
 *
 * Matrices are ELL: entry j of row i at cols[j*ld + i] / vals[j*ld + i],
 * width <= 32, padding entries with value 0 and a valid column.
 * The operator (cols/vals) is given in SOLVER FORM, built once per matrix by
 * smx_pcg_prepare from A and dinv = 1/diag(A): the diagonal entry of every row
 * in slot 0 and every value scaled by dinv of its column (A D^-1).  The call
 * fails with SMX_ERR_VALUE, before any iteration, when some row's slot 0 is
 * not its diagonal.  The raw operator of mode 1 stays plain A_raw.
 * mode 0 (u1, Neumann): solves A x = b - mean(b)*project, where mean(b) is
 *     (sum of the nb_part device partial sums b_part) / n; x0 = x if warm
 *     else 0.
 * mode 1 (u2, Dirichlet lifting): A = A_bc; boundary rows take the data
 *     x_k = gb[bpos[k]] where bpos[k] >= 0; interior rows solve
 *     sum_j A_raw_kj x_j = 0 (raw stiffness in ELL, raw_width <= 32).
 * nd (trailing identity rows, 0 for none): the system has n + nd rows, and
 *     rows n .. n+nd-1 of the solver order are identity rows (the Dirichlet
 *     rows of A_bc): their x is their data (mode 1: gb[bpos[row]]; mode 0:
 *     b of that row, which needs project = 0), their residual is 0, and they
 *     count in ||b|| as rows of the full system, so the tolerance is the full
 *     system's.  No iteration touches them: cols/vals describe rows 0..n-1
 *     only and must reference columns < n only (A_bc's interior block); the
 *     raw operator of mode 1 (rows 0..n-1) may reference every column, and
 *     bpos, perm, b and x cover n + nd rows.  An iteration then streams the
 *     n interior rows alone.  With nd = 0 Dirichlet rows may also sit among
 *     the n rows (mode 1 treats any row with bpos >= 0 as one).
 * perm (optional, NULL = identity): solver row i is caller row perm[i].  The
 *     matrices (cols/vals, raw), dinv and bpos are then given in solver order
 *     (P A P^T etc.); b, gb and x stay in the caller's order (b is gathered
 *     on entry, x scattered back on exit).  A bandwidth-reducing order keeps
 *     the SpMV neighbours of each XCD's contiguous row range in its own L2.
 * Convergence: ||r||_2 <= max(rtol*||b||_2, atol), b the (lifted) right side.
 * work: smx_pcg_ws_size(n + nd) doubles.  On return work[SMX_PCG_SUMX] holds
 * sum(x) over the n iterated rows on the device (the nullspace projection of x, read by later kernels).
 * its_out / rnorm_out / sums_out: host pointers filled on return (the call
 * synchronizes `stream` every `check_every` iterations; the first wait comes
 * after expected_its + 1 iterations when expected_its > 0).
 * Speculative launch (check_every in [-64, -1], expected_its > 0): the call
 * enqueues expected_its + 1 - check_every iterations and returns at once,
 * without waiting (its_out = -1, rnorm_out and sums_out untouched).
 * Iterations past convergence return at once on the device's done flag, so
 * the result equals the waiting call's whenever the solve converged within
 * them.  The status block at the start of work then tells, once the stream
 * has reached it: work[SMX_PCG_DONE] != 0 when the solve stopped,
 * work[SMX_PCG_ITERS] its iteration count, work[SMX_PCG_FAIL] 0 (converged),
 * 1 (breakdown), 2 (max_it) or 3 (operator not in solver form).  A caller
 * that finds DONE == 0 must solve again (waiting) before using x.  The same
 * holds for smx_pcg_amg.
 * ---------------------------------------------------------------------- */
#define SMX_PCG_DONE 0
#define SMX_PCG_ITERS 1
#define SMX_PCG_FAIL 5
#define SMX_PCG_SUMX 6
int64_t smx_pcg_ws_size(int64_t n);
/* Solver form of a CG operator (replaces nothing in the reference: PETSc's
 * KSP takes A itself; this is the one-time set-up that a PCSetUp would do).
 * Per row: the first slot whose column is the row moves to slot 0 (slot 0's
 * entry takes its place) and every value becomes a_ij * dinv[col].  A row
 * without a diagonal entry is copied unchanged.  cols_out/vals_out may alias
 * cols/vals (in-place).  Applying it twice scales twice: call it once. */
int smx_pcg_prepare(int64_t n, int32_t width, int64_t ld, const int32_t* cols,
                    const double* vals, const double* dinv, int32_t* cols_out,
                    double* vals_out, void* stream);
int smx_pcg_jacobi(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols,
                   const double* vals, const double* dinv, int32_t mode,
                   const double* b, const double* b_part, int64_t nb_part,
                   int32_t project, int32_t raw_width, int64_t raw_ld,
                   const int32_t* raw_cols, const double* raw_vals,
                   const int32_t* bpos, const double* gb, double* x, int32_t warm,
                   const int32_t* perm, double rtol, double atol, int32_t max_it,
                   int32_t check_every, int32_t expected_its, double* work,
                   int32_t* its_out, double* rnorm_out, double* sums_out, void* stream);

/* Mixed-precision Poisson solve: FP32 inner CG, FP64 iterative refinement.
 * The same solve as smx_pcg_jacobi / smx_pcg_amg (same operators, tolerance
 * test, modes, workspace and status block), with
 *   - the iteration vectors (r, w, s, p, and u for the AMG form) in FP32,
 *     and the iteration operator read from `vals32`: FP32 values in the
 *     layout of `vals` (the solver form for the Jacobi solver, the plain A
 *     for the AMG solver);
 *   - x, the CG scalars and the inner products in FP64: each product of two
 *     FP32 values is exact in FP64, the wave sums are FP64 and the sums
 *     across waves and workgroups are compensated (TwoSum);
 *   - iterative refinement: the FP32 solve runs to rtol_inner (below rtol
 *     leaves room for the FP32 drift of the recurrence residual; above it
 *     stages the solve); then, up to `refinements` times, the residual b - A x is recomputed
 *     from the FP64 operator (`vals`) and checked against rtol; when it
 *     fails, a correction solve (FP32 again, from the current x) runs to
 *     max(rtol, rtol_inner^(k+1)) of ||b||, the last one to rtol.
 *     refinements = 0: the FP32 solve alone.
 * A speculative launch (check_every < 0) enqueues the FP32 solve
 * (expected_its + 1 - check_every iterations) and one FP64 check; a failed
 * check leaves SMX_PCG_DONE = 0 and the caller solves again waiting.  The
 * status block holds the last solve's iteration count at SMX_PCG_ITERS and
 * the earlier solves' total at SMX_PCG_ITERS0; its_out (waiting form) is
 * their sum.  Mixed-precision form of the reference's KSP cg solves
 * (Demag_Lindholm_GPU.py:216-226, 271-281, 569-640), which run in FP64. */
typedef struct {
  const float* vals32;      /* FP32 iteration operator values */
  double rtol_inner;        /* tolerance of the first FP32 solve; powers of it stage the corrections */
  int32_t refinements;      /* FP64 residual checks / correction solves, 0..8 */
  int32_t pad;
} smx_cg_mixed;
#define SMX_PCG_ITERS0 7
int smx_pcg_jacobi_mixed(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols,
                         const double* vals, const double* dinv, int32_t mode,
                         const double* b, const double* b_part, int64_t nb_part,
                         int32_t project, int32_t raw_width, int64_t raw_ld,
                         const int32_t* raw_cols, const double* raw_vals,
                         const int32_t* bpos, const double* gb, double* x, int32_t warm,
                         const int32_t* perm, double rtol, double atol, int32_t max_it,
                         int32_t check_every, int32_t expected_its, double* work,
                         int32_t* its_out, double* rnorm_out, double* sums_out,
                         const smx_cg_mixed* mixed, void* stream);

/* ---------------------------------------------------------------------- *
 * (c') CG preconditioned by one smoothed-aggregation AMG V-cycle.  Replaces
 * PETSc KSP cg + PC gamg of the two Poisson solves (Demag_Lindholm_GPU.py:
 * 216-226 for u1, 271-281 for u2; solves at :569-582, :618-640), whose
 * defaults are one V-cycle per application with Jacobi-type smoothing.
 *
 * The hierarchy (built on the host) gives, per level l with n_l unknowns,
 *   S_l (n_l x n_l)      = 2w D^-1 - w^2 D^-1 A_l D^-1
 *   Q_l (n_l x n_{l+1})  = (I - w D^-1 A_l) P_l     (P_l smoothed prolongator)
 *   R_l (n_{l+1} x n_l)  = Q_l^T
 * and on the last level only S = a dense (pseudo-)inverse of the coarse
 * operator.  The preconditioner is M^-1 b = S_0 b + Q_0 M_1^-1 (R_0 b),
 * recursively: one damped-Jacobi pre- and post-smoothing sweep per level.
 *
 * Operators are sliced ELL for `lanes` lanes per row (1..64, a power of two:
 * s_tpr for S and Q of a level, r_tpr for R): slice s holds rows
 * s*rps .. s*rps+rps-1 with rps = 256/lanes (the last slice padded to rps
 * rows), all of width W_s = (slice_off[s+1]-slice_off[s])/rps; slot j of
 * local row i sits at slice_off[s] + j*rps + i when lanes == 1 and at
 * slice_off[s] + i*W_s + j otherwise.  Padding slots hold value 0 and a valid
 * column.  col == NULL: dense rows, the column of slot j is j.  Values are
 * FP64 or FP32 (val_dtype, the same for every operator); offsets and columns
 * int32.  Levels >= 1 bring device vectors b and x of n_l doubles.
 *
 * Everything else (ELL operator = plain A in solver order, modes, perm,
 * convergence test, status record, SMX_PCG_SUMX) is as smx_pcg_jacobi.
 * work: smx_pcg_amg_ws_size(n + nd) doubles; levels[0].n == n.
 * ---------------------------------------------------------------------- */
typedef struct {
  const int32_t* slice_off;  /* nslices + 1 slot offsets */
  const int32_t* col;        /* NULL: dense rows */
  const void* val;
  int64_t rows;
} smx_amg_op;

typedef struct {
  int64_t n;
  smx_amg_op S, Q, R;
  int32_t s_tpr, r_tpr;
  double* b;
  double* x;
} smx_amg_level;

int64_t smx_pcg_amg_ws_size(int64_t n);
/* u = M^-1 r: one application of the V-cycle above (PETSc PCApply of the
 * GAMG preconditioner).  r and u hold levels[0].n doubles; work holds at
 * least smx_pcg_amg_ws_size(levels[0].n) doubles. */
int smx_amg_apply(const smx_amg_level* levels, int32_t nlev, int32_t val_dtype, const double* r,
                  double* u, double* work, void* stream);
int smx_pcg_amg(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols, const double* vals,
                int32_t mode, const double* b, const double* b_part, int64_t nb_part,
                int32_t project, int32_t raw_width, int64_t raw_ld, const int32_t* raw_cols,
                const double* raw_vals, const int32_t* bpos, const double* gb, double* x,
                int32_t warm, const int32_t* perm, const smx_amg_level* levels, int32_t nlev,
                int32_t val_dtype, double rtol, double atol, int32_t max_it,
                int32_t check_every, int32_t expected_its, double* work, int32_t* its_out,
                double* rnorm_out, double* sums_out, void* stream);
/* Mixed-precision form of smx_pcg_amg (FP32 inner CG with the FP32 values of
 * A in mixed->vals32, FP64 refinement), as smx_pcg_jacobi_mixed describes. */
int smx_pcg_amg_mixed(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols,
                      const double* vals, int32_t mode, const double* b, const double* b_part,
                      int64_t nb_part, int32_t project, int32_t raw_width, int64_t raw_ld,
                      const int32_t* raw_cols, const double* raw_vals, const int32_t* bpos,
                      const double* gb, double* x, int32_t warm, const int32_t* perm,
                      const smx_amg_level* levels, int32_t nlev, int32_t val_dtype, double rtol,
                      double atol, int32_t max_it, int32_t check_every, int32_t expected_its,
                      double* work, int32_t* its_out, double* rnorm_out, double* sums_out,
                      const smx_cg_mixed* mixed, void* stream);

/* ---------------------------------------------------------------------- *
 * (d) Fused effective field + LLG right-hand side.
 * Replaces apply_linear_field_vec + anisotropy + demag G.mult/V + axpys +
 * _LLG_RHS_KERNEL (llg_module_GPU.py:100-122, 962-1006, 1261-1285;
 * Anisotropy_GPU.py:25-38; Demag_Lindholm_GPU.py:660-670).  One thread per
 * node over the node-graph SELL-64 pattern:
 *   H = ex*m + dmi*m + pref_ani (m.n) n + [grad*(u1+u2-u1_sum/n)]/V + H0
 *       + pref_cub (s1 c1 + s2 c2 + s3 c3)            (cubic anisotropy)
 *   rhs = P (p m x H + alpha m x (m x H)) + S (1 - |m|^2) m
 *       - stt_pref (stt_a m x Z + stt_b m x (m x Z))   (Zhang-Li STT, Z = K_J m)
 *       + sot_fl m x p + sot_dl m x (m x p)             (spin-orbit torque)
 * with a_i = m.c_i, s1 = a1 (a2^2 + a3^2) (cyclic) for the orthonormal cubic
 * axes c_i (Cubic_Anisotropy_GPU.py:33-53, 205-276), K_J the Zhang-Li
 * operator (J.grad) with rows scaled by 1/V_i (llg_stt_module_GPU.py:393-435,
 * 101-139) and the spin-orbit coefficients of llg_SOT_module_GPU.py:170-205,
 * 813-817.  Any operator pointer may be NULL (term absent); a zero
 * prefactor disables its term.  Outputs may be NULL.
 * ---------------------------------------------------------------------- */
typedef struct {
  int64_t n;                 /* nodes */
  const int64_t* slice_ptr;  /* node graph (shared by ex/dmi/grad) */
  const int32_t* cols;
  const double* ex_vals;     /* k=1 exchange (prefactor folded), or NULL */
  const double* dmi_vals;    /* DMI blocks (prefactor folded; layout: dmi_ncomp), or NULL */
  const double* grad_vals;   /* k=3 demag gradient G, or NULL */
  const double* u1;          /* N */
  const double* u2;          /* N */
  const double* u1_sum;      /* device scalar: sum(u1) (projected out), or NULL */
  const double* inv_vol;     /* N: 1/V_i */
  const double* m;           /* 3N */
  const double* ani_axis;    /* 3N unit axes, or NULL -> ani_axis_u */
  double ani_axis_u[3];
  double ani_pref;           /* 2 Ku/(mu0 Ms); 0 disables */
  const double* h0;          /* 3N external field, or NULL -> h0_u */
  double h0_u[3];
  double* h_eff;             /* 3N out (total field) or NULL */
  double* h_demag;           /* 3N out (demag part) or NULL */
  double* rhs;               /* 3N out or NULL */
  double P;                  /* -gamma/(1+alpha^2) */
  double alpha;
  double do_precess;
  double stab;               /* Ms gamma/(1+alpha^2)/2 */
  double cub_pref;           /* 2 Kc1/(mu0 Ms); 0 disables */
  const double* cub_axes;    /* 9N: c1, c2, c3 of every node, or NULL -> cub_axes_u */
  double cub_axes_u[9];
  const double* stt_vals;    /* k=1 Zhang-Li operator on the node graph, or NULL */
  double stt_pref;           /* J P muB / (e Ms (1+beta^2)) / (1+alpha^2) / 1e-9 */
  double stt_a;              /* beta - alpha */
  double stt_b;              /* 1 + alpha beta */
  const double* sot_p;       /* 3N spin polarization, or NULL -> sot_p_u */
  double sot_p_u[3];
  double sot_fl;             /* -gamma (H_fl - alpha H_dl)/(1+alpha^2) */
  double sot_dl;             /* -gamma (H_dl + alpha H_fl)/(1+alpha^2) */
  /* DMI value layout: dmi_ncomp values per graph slot (k = dmi_ncomp in the
   * SELL convention); component q = 3 i + j of the 3x3 block (field
   * component i, m component j) at dmi_off[q], absent (exactly zero) when
   * dmi_off[q] < 0.  dmi_ncomp == 0: all nine, in order.  Interfacial DMI
   * with n0 along z needs 5 (xz, yz, zx, zy, zz), bulk DMI 6 (no diagonal).
   * Either layout gives the same bits: a row's present components are taken
   * in column order, the first as a product and the rest by fma. */
  int32_t dmi_ncomp;
  int32_t dmi_off[9];
} smx_field_args;

int smx_field_llg_rhs(const smx_field_args* a, void* stream);

/* Jacobian-vector product of the LLG RHS with the field derivative
 * Hv = ex*v + dmi*v + pref_ani (v.n) n + dH_cub/dm v computed in-kernel (demag
 * and external field excluded, as compute_H_jac_vec / apply_jacobian_field_vec
 * do), plus the exact derivatives of the STT and SOT torques.  Replaces
 * _JAC_VEC_KERNEL + jac_vec_times_vec (llg_module_GPU.py:125-164, 860-883,
 * 1293-1323), _JAC_VEC_STT_KERNEL (llg_stt_module_GPU.py:142-215) and
 * _JAC_VEC_SOT_KERNEL (llg_SOT_module_GPU.py:208-275).  `a->m` is the
 * Jacobian state, `hm` its field H_jac(m). */
int smx_llg_jvp(const smx_field_args* a, const double* hm, const double* v,
                double* out, double gamma, void* stream);

/* ---------------------------------------------------------------------- *
 * (f) Implicit (BDF) LLG steps.  smx_llg_pc_build replaces
 * JvContext.update_pc_full_fast_gpu + _PC_BUILD_INV3_KERNEL
 * (llg_module_GPU.py:165-290, 1447-1505): inv (9N, row-major 3x3 per node) =
 * [(shift + eps_reg) I - J_i]^-1 with the local block
 * J_i = -+ c1 d(m x H)/dm -+ c2 d(m x (m x H))/dm + S ((1-|m|^2) I - 2 m m^T)
 * (c1 = gamma p/(1+alpha^2), c2 = gamma alpha/(1+alpha^2); the derivatives
 * taken with dH/dm = kappa I, kappa = mean (use_abs_kappa: mean |.|) of the
 * diagonal of the linear field operator of `a`: exchange, DMI, uniaxial and
 * cubic anisotropy) at the state a->m with field hm; upper signs (exact = 1):
 * the derivative of the RHS; lower signs (exact = 0): the reference's block;
 * the stabilization term when include_stab; |det| < det_eps is pushed to
 * det_eps.  smx_block3_apply replaces JvContext.apply
 * (_PC_APPLY_KERNEL, :293-307): y_i = inv_i x_i.
 * smx_multi_dot: out[j] = V_j . w for j < k (V_j at V + j*ldv, len doubles;
 * deterministic two-stage reduction; work: smx_multi_dot_ws_size doubles).
 * smx_lincomb: out = beta out + sum_j c_j V_j (c: k device doubles).
 * These carry the Arnoldi steps of the GMRES that replaces PETSc KSP gmres
 * (:2074-2115). */
int smx_llg_pc_build(const smx_field_args* a, const double* hm, double shift, double eps_reg,
                     double det_eps, int32_t include_stab, int32_t use_abs_kappa, int32_t exact,
                     double* inv, void* stream);
int smx_block3_apply(int64_t n, const double* inv, const double* x, double* y, void* stream);
int64_t smx_multi_dot_ws_size(int64_t len, int32_t k);
int smx_multi_dot(int64_t len, int32_t k, const double* V, int64_t ldv, const double* w,
                  double* out, double* work, void* stream);
int smx_lincomb(int64_t len, int32_t k, const double* V, int64_t ldv, const double* c,
                double beta, double* out, void* stream);

/* Energy minimizer steps (Minimizer_GPU.py:46-88, 306-416): the projected
 * direction d = h - (m.h) m, and the norm-preserving Cayley update
 * out = (I - C)^-1 (I + C) m with C v = c x v, c = (step/2) (m x d) (a rotation
 * of m about c by 2 atan|c|; out = m + step d + O(step^2) for d tangent). */
int smx_projected_direction(int64_t n, const double* m, const double* h, double* d, void* stream);
int smx_cayley_step(int64_t n, const double* m, const double* d, double step, double* out,
                    void* stream);
/* m_i <- m_i / |m_i| (Minimizer_GPU.py:238-250). */
int smx_normalize3(int64_t n, double* m, void* stream);
/* out = max_i |v_i| over 3-vectors (deterministic); red as smx_sell_spmv. */
int smx_max_norm3(int64_t n, const double* v, double* out, double* red, void* stream);

/* out = sum_i w_i (a_i . b_i) over 3-vectors (w may be NULL -> 1); `red` as in
 * smx_sell_spmv.  Energies (Demag_Lindholm_GPU.py:706-724,
 * llg_module_GPU.py:1014-1051). */
int smx_dot3_weighted(int64_t n, const double* a, const double* b,
                      const double* w, double* out, double* red, void* stream);

/* Dormand-Prince stage combination: out = y + dt * sum_s c[s] k_s over
 * nstage (<=7) stage vectors (host array of device pointers), length len.
 * When err_out != NULL, also err = dt*sum_s e[s] k_s and the weighted RMS
 * norm sqrt(mean((err/(atol+rtol*max(|y|,|out|)))^2)) into *err_out. */
int smx_rk_combine(int64_t len, const double* y, double dt, int32_t nstage,
                   const double* const* k, const double* c, const double* e,
                   double atol, double rtol, double* out, double* err_out,
                   double* red, void* stream);

/* ---------------------------------------------------------------------- *
 * (a22) Energy with per-cell Ms and A and its adjoint (no counterpart in the
 * reference, which has no adjoint; the energies reduce to its lumped
 * exchange / anisotropy / demag energies, Exchange_GPU.py:85-92,
 * Anisotropy_GPU.py:250-275, Demag_Lindholm_GPU.py:706-724, for uniform Ms, A).
 * Cell data: cells (T x 4 int32), ck = vol grad(l_a).grad(l_b) (T x 16),
 * cgv = vol/4 grad(l_a) (T x 4 x 3).  inc_ptr/inc: node -> 4*cell + local
 * vertex incidence.  `partials` receive one sum per 256-entry block
 * (smx_energy_blocks(n) of them).
 *   cell_forward:  s_c = Ms_c sum_b m_b;  f_{c,a} = 2e-9 A_c sum_b K_ab m_b;
 *                  e_c = 1e-9 m_c^T K m_c (= dE/dA_c); partials of A_c e_c
 *   node_div:      b_a = sum_{(c,l) at a} cgv[c,l] . s_c      (Div_Ms m)
 *   cell_backward: t_c = Ms_c sum_a y_a cgv[c,a]; dMs_c = (sum_a y_a cgv[c,a]).sum_b m_b
 *   node_grad:     grad_a = sum (f[c,l] + t_c) - 2e-27 Ku V_a (m_a.n) n; partials of E_an
 *   vec_glue:      op 0 interior restriction, 1 boundary residual x[idx]-y[idx],
 *                  2 boundary scatter-add, 3 a(x - sx/nrm) + b(y - sy/nrm)
 * ---------------------------------------------------------------------- */
int64_t smx_energy_blocks(int64_t n);
int smx_energy_cell_forward(int64_t T, const int32_t* cells, const double* ck, const double* Ms,
                            const double* A, const double* m, double* s, double* f, double* e,
                            double* partials, void* stream);
int smx_energy_node_div(int64_t N, const int64_t* inc_ptr, const int32_t* inc, const double* cgv,
                        const double* s, double* b, double* partials, void* stream);
int smx_energy_cell_backward(int64_t T, const int32_t* cells, const double* cgv, const double* Ms,
                             const double* y, const double* m, double* t, double* dMs,
                             void* stream);
int smx_energy_node_grad(int64_t N, const int64_t* inc_ptr, const int32_t* inc, const double* f,
                         const double* t, const double* m, const double* vol, const double* axis,
                         const double* axis_u, double Ku, double* grad, double* partials,
                         void* stream);
int smx_vec_glue(int32_t op, int64_t n, const double* x, const double* y, const int32_t* idx,
                 double a, double b, const double* sx, const double* sy, double nrm, double* out,
                 void* stream);
int smx_dot_partials(int64_t n, const double* x, const double* y, double* partials, void* stream);

/* ---------------------------------------------------------------------- *
 * (a22) LLG right-hand side with per-cell Ms and its vector-Jacobian product,
 * for the adjoint through the Dormand-Prince steps (no counterpart in the
 * reference; the forward is _LLG_RHS_KERNEL, llg_module_GPU.py:100-122, with
 * a nodal Ms_i = sum_c |c|/4 Ms_c / V_i in the exchange prefactor and in S).
 *   node_rhs:     r = P (p m x H + alpha m x (m x H)) + S_i (1 - |m|^2) m
 *   node_rhs_vjp: for the cotangent lam of r: gm (explicit m part), gk = cex gH,
 *                 gd = gH / V, gnode = gH.(-Hex/Ms_i) + s0 (1-|m|^2)(m.lam)
 *                 with gH = dr/dH^T lam and Hex = H - Hd - H0
 *   node_to_cell: gcell_c (+)= |c|/4 sum_{i in c} gnode_i / V_i
 *   cell_to_node: Ms_i = sum_{c ni i} |c|/4 Ms_c / V_i   (qvol = |c|/4)
 *   row_scale:    out[e] = vals[e] * scale[rows[e]] (rows < 0: out 0)
 * ---------------------------------------------------------------------- */
int smx_node_rhs(int64_t n, const double* m, const double* h, const double* S, double P,
                 double alpha, double do_precess, double* r, void* stream);
int smx_node_rhs_vjp(int64_t n, const double* m, const double* h, const double* hd,
                     const double* h0_u, const double* Ms, const double* S, double s0,
                     const double* cex, const double* inv_vol, double P, double alpha,
                     double do_precess, const double* lam, double* gm, double* gk,
                     double* gd, double* gnode, void* stream);
int smx_node_to_cell(int64_t T, const int32_t* cells, const double* qvol, const double* inv_vol,
                     const double* gnode, int32_t accumulate, double* gcell, void* stream);
int smx_cell_to_node(int64_t N, const int64_t* inc_ptr, const int32_t* inc, const double* qvol,
                     const double* inv_vol, const double* Ms_cell, double* Ms_node, void* stream);
int smx_row_scale(int64_t nslots, const double* vals, const int32_t* rows, const double* scale,
                  double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SMX_HIP_H */
