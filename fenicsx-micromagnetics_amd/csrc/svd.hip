// Truncated SVD and randomized-SVD compression of batches of dense H-matrix
// blocks, one 1024-thread workgroup per block, no library calls.
//
// Truncated SVD (smx_svd_batched; the semantics of truncated_svd_gpu,
// lindholm_hmatrix_gpu/compressors.py:373-414):
//   1. Householder QR with column pivoting of X = A^T (the rows of the
//      row-major block are the columns of X, contiguous), pivoting on the
//      largest remaining column norm (first maximum), stopped once the
//      trailing Frobenius norm is <= kFloorHi * rtol * ||A||_F (or, when
//      kJMax steps do not get there, accepted at kFloorLo * rtol).  Column
//      norms are recomputed exactly in every update pass (no downdating).
//   2. One-sided Jacobi SVD of the k leading rows T of the triangular factor.
//   Since X P = Q_H [T; E] with Q_H orthogonal, A^T A = P (T^T T + E^T E) P^T:
//   every captured energy sum_{i<=r} s_i^2 of T is within ||E||_F^2 of A's,
//   so the rank rule sees A's spectrum to (kFloor rtol)^2 ||A||^2 (1e-8 or
//   1e-4 of rtol^2), the rank-r factors are within about ||E||_F of A's
//   truncated SVD, and ||A - U V||_F^2 = ||A||^2 - sum_{i<=r} s_i(T)^2
//   exactly (the reported residual is the stored factors' residual).  A block
//   that does not reach kFloorLo within kJMax QR steps is reported as failed
//   (status 4), so the builder subdivides it.
// Randomized SVD (smx_rsvd_round; one round of randomized_svd_gpu,
// compressors.py:417-492): Y = A Omega, `power` times Q = orth(Y),
// Y = A (A^T Q); Q = orth(Y); B = Q^T A; SVD of B.  orth() is the same
// one-sided Jacobi on the rows of Y^T, each nonzero row normalized (Jacobi's
// relative stopping rule leaves even rows at round-off level orthogonal to
// the others, so the normalized set is orthonormal, as a Householder Q).
// The host draws Omega per block and repeats rounds with a doubled target
// for blocks whose sketch found no rank meeting rtol (status 7).
//
// One-sided Jacobi: plane rotations of row pairs in tournament order (k-1
// rounds of k/2 disjoint pairs, one wavefront per pair), rotation from
// zeta = (b - a) / 2g, t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2))
// (Rutishauser), until every pair has |g| <= sqrt(L) eps sqrt(a b).  The
// rotations are accumulated in J (J R_in = R_out).  Every sum runs in a fixed
// order, so the factors are deterministic.
//
// Rank rule (compressors.py:402-406, 470-476): with s sorted descending,
// captured = cumsum(s^2), rel_r = sqrt(max(||A||^2 - captured_r, 0)) /
// sqrt(||A||^2); the rank is the first r with rel_r <= rtol, else the
// usable maximum.  Factors: column r of U (scaled by s_r) at u_off + r m,
// row r of V at v_off + r n, as the ACA kernel stores them.
#include "smx_common.h"

namespace {

constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;
constexpr int kJMax = SMX_SVD_KMAX;
constexpr int kMaxSweeps = 40;
constexpr double kFloorHi = 1e-4;
constexpr double kFloorLo = 1e-2;
constexpr double kEps = 1.1102230246251565e-16;  // 2^-53
constexpr int kChunk = 8;                        // sketch rows per GEMM pass

enum { ST_STORED = 0, ST_OVER_CAP = 2, ST_ZERO = 3, ST_FAILED = 4, ST_WIDER = 7 };

struct Shared {
  double red[kWaves];
  double redv[kWaves];
  int redi[kWaves];
  double sig[kJMax];
  int ord[kJMax];
  double bcast[4];
  int flag;
  int ibcast[4];
};

__device__ __forceinline__ int wave_id() { return (int)threadIdx.x >> 6; }
__device__ __forceinline__ int lane_id() { return (int)threadIdx.x & 63; }

// Sum over the workgroup in a fixed order (wave totals by DPP, then the
// waves' totals in wave order); the result is returned to every thread.
__device__ double block_sum(double v, Shared& sh) {
  v = smx::wave_total(v);
  if (lane_id() == 0) sh.red[wave_id()] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) s += sh.red[w];
  __syncthreads();
  return s;
}

// First position of the maximum (ties to the lower index) over the
// workgroup; threads without candidates pass (-1, INT_MAX).
__device__ void block_argmax(double& v, int& i, Shared& sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(i, off, 64);
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
  if (lane_id() == 0) {
    sh.redv[wave_id()] = v;
    sh.redi[wave_id()] = i;
  }
  __syncthreads();
  v = sh.redv[0];
  i = sh.redi[0];
#pragma unroll
  for (int w = 1; w < kWaves; ++w) {
    const double ov = sh.redv[w];
    const int oi = sh.redi[w];
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
  __syncthreads();
}

// ||A||_F^2 of the row-major m x n block, fixed order.
__device__ double block_norm_sq(const double* __restrict__ Ab, int m, int n, Shared& sh) {
  double s = 0.0;
  const int64_t mn = (int64_t)m * n;
  for (int64_t e = threadIdx.x; e < mn; e += kThreads) s = fma(Ab[e], Ab[e], s);
  return block_sum(s, sh);
}

// Plane rotation of rows p and q (length L) of R, and of rows p and q
// (length k) of J, by one wavefront; sets *flag when it rotates.
__device__ void jacobi_pair(double* __restrict__ R, int64_t ld, int p, int q, int L, double* J, int k,
                            double tol, int* flag) {
  const int lane = lane_id();
  double* rp = R + (int64_t)p * ld;
  double* rq = R + (int64_t)q * ld;
  double a = 0.0, b = 0.0, g = 0.0;
  for (int l = lane; l < L; l += 64) {
    const double x = rp[l], y = rq[l];
    a = fma(x, x, a);
    b = fma(y, y, b);
    g = fma(x, y, g);
  }
  a = smx::wave_total(a);
  b = smx::wave_total(b);
  g = smx::wave_total(g);
  if (!(fabs(g) > tol * sqrt(a) * sqrt(b))) return;  // uniform
  const double zeta = (b - a) / (2.0 * g);
  const double t = fabs(zeta) > 1e150 ? 0.5 / zeta
                                      : copysign(1.0, zeta) / (fabs(zeta) + sqrt(fma(zeta, zeta, 1.0)));
  const double c = 1.0 / sqrt(fma(t, t, 1.0));
  const double s = c * t;
  for (int l = lane; l < L; l += 64) {
    const double x = rp[l], y = rq[l];
    rp[l] = c * x - s * y;
    rq[l] = s * x + c * y;
  }
  if (J) {
    double* jp = J + (int64_t)p * k;
    double* jq = J + (int64_t)q * k;
    for (int l = lane; l < k; l += 64) {
      const double x = jp[l], y = jq[l];
      jp[l] = c * x - s * y;
      jq[l] = s * x + c * y;
    }
  }
  if (lane == 0) *flag = 1;
}

// Tournament position of player x in round r (kp players, kp even).
__device__ __forceinline__ int tour(int x, int r, int kp) {
  return x == 0 ? 0 : 1 + (x - 1 + r) % (kp - 1);
}

// One-sided Jacobi on the k rows (length L, leading dimension ld) of R; J
// (k x k, row-major), when given, must hold the identity on entry.
__device__ void jacobi_rows(double* R, int64_t ld, int k, int L, double* J, Shared& sh) {
  const int kp = k + (k & 1);
  const int npairs = kp / 2;
  const double tol = kEps * sqrt((double)(L > 1 ? L : 1));
  for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
    if (threadIdx.x == 0) sh.flag = 0;
    __syncthreads();
    for (int r = 0; r + 1 < kp; ++r) {
      for (int a = wave_id(); a < npairs; a += kWaves) {
        int p = tour(a, r, kp), q = tour(kp - 1 - a, r, kp);
        if (p > q) {
          const int tmp = p;
          p = q;
          q = tmp;
        }
        if (q < k) jacobi_pair(R, ld, p, q, L, J, k, tol, &sh.flag);
      }
      __syncthreads();
    }
    const bool again = sh.flag != 0;
    __syncthreads();
    if (!again) break;
  }
}

// Row norms of R (k rows) into sh.sig and the descending order (ties to the
// lower row) into sh.ord.
__device__ void row_norms_sorted(const double* __restrict__ R, int64_t ld, int k, int L, Shared& sh) {
  for (int i = wave_id(); i < k; i += kWaves) {
    const double* ri = R + (int64_t)i * ld;
    double s = 0.0;
    for (int l = lane_id(); l < L; l += 64) s = fma(ri[l], ri[l], s);
    s = smx::wave_total(s);
    if (lane_id() == 0) sh.sig[i] = sqrt(s);
  }
  __syncthreads();
  if ((int)threadIdx.x < k) {
    const int i = threadIdx.x;
    const double si = sh.sig[i];
    int r = 0;
    for (int j = 0; j < k; ++j) {
      const double sj = sh.sig[j];
      r += (sj > si || (sj == si && j < i)) ? 1 : 0;
    }
    sh.ord[r] = i;
  }
  __syncthreads();
}

// Rank rule on the sorted values sh.sig[sh.ord[0..]] (thread 0): scans
// r = 1 .. min(usable, nv).  Returns found (some r met rtol); rank and rel as
// the reference's rule when found or when usable <= nv.
__device__ bool rank_rule(const Shared& sh, int nv, int usable, double norm_sq, double rtol, int& rank,
                          double& rel) {
  const double nA = sqrt(norm_sq);
  double captured = 0.0;
  const int lim = usable < nv ? usable : nv;
  rank = lim;
  rel = 1.0;
  for (int r = 0; r < lim; ++r) {
    const double s = sh.sig[sh.ord[r]];
    captured += s * s;
    const double v = sqrt(fmax(norm_sq - captured, 0.0)) / nA;
    rel = v;
    if (v <= rtol) {
      rank = r + 1;
      return true;
    }
  }
  return false;
}

// C[i, row] = sum_l M(i, l) A[row, l] for i < s, row < m (A row-major m x n,
// M(i, l) = M[i * msi + l * msl]); one wavefront per row of A, kChunk sketch
// rows per pass.  C is row-major s x m.
__device__ void gemm_at(const double* __restrict__ M, int64_t msi, int64_t msl, const double* __restrict__ Ab,
                        int m, int n, int s, double* __restrict__ C) {
  const int lane = lane_id();
  for (int row = wave_id(); row < m; row += kWaves) {
    const double* ar = Ab + (int64_t)row * n;
    for (int i0 = 0; i0 < s; i0 += kChunk) {
      double acc[kChunk];
#pragma unroll
      for (int ii = 0; ii < kChunk; ++ii) acc[ii] = 0.0;
      for (int l = lane; l < n; l += 64) {
        const double a = ar[l];
#pragma unroll
        for (int ii = 0; ii < kChunk; ++ii) {
          if (i0 + ii < s) acc[ii] = fma(M[(int64_t)(i0 + ii) * msi + (int64_t)l * msl], a, acc[ii]);
        }
      }
#pragma unroll
      for (int ii = 0; ii < kChunk; ++ii) {
        const double v = smx::wave_total(acc[ii]);
        if (lane == 0 && i0 + ii < s) C[(int64_t)(i0 + ii) * m + row] = v;
      }
    }
  }
}

// C[i, j] = sum_row M[i, row] A[row, j] for i < s, j < n (M row-major s x m);
// one thread per column j, kChunk sketch rows per pass.  C is row-major s x n.
__device__ void gemm_a(const double* __restrict__ M, const double* __restrict__ Ab, int m, int n, int s,
                       double* __restrict__ C) {
  for (int j = threadIdx.x; j < n; j += kThreads) {
    for (int i0 = 0; i0 < s; i0 += kChunk) {
      double acc[kChunk];
#pragma unroll
      for (int ii = 0; ii < kChunk; ++ii) acc[ii] = 0.0;
      for (int row = 0; row < m; ++row) {
        const double a = Ab[(int64_t)row * n + j];
#pragma unroll
        for (int ii = 0; ii < kChunk; ++ii) {
          if (i0 + ii < s) acc[ii] = fma(M[(int64_t)(i0 + ii) * m + row], a, acc[ii]);
        }
      }
#pragma unroll
      for (int ii = 0; ii < kChunk; ++ii) {
        if (i0 + ii < s) C[(int64_t)(i0 + ii) * n + j] = acc[ii];
      }
    }
  }
}

// Orthonormal rows spanning the rows of Y (s x m): Jacobi, then every
// nonzero row scaled to unit length.
__device__ void orth_rows(double* Y, int s, int m, Shared& sh) {
  jacobi_rows(Y, m, s, m, nullptr, sh);
  for (int i = wave_id(); i < s; i += kWaves) {
    double* yi = Y + (int64_t)i * m;
    double q = 0.0;
    for (int l = lane_id(); l < m; l += 64) q = fma(yi[l], yi[l], q);
    q = smx::wave_total(q);
    const double inv = q > 0.0 ? 1.0 / sqrt(q) : 0.0;
    for (int l = lane_id(); l < m; l += 64) yi[l] *= inv;
  }
  __syncthreads();
}

__device__ void set_identity(double* J, int k) {
  for (int e = threadIdx.x; e < k * k; e += kThreads) J[e] = (e / k == e % k) ? 1.0 : 0.0;
  __syncthreads();
}

__device__ void finish_status(int32_t* rank_out, double* resid_out, int32_t* status_out, int rank, double rel,
                              int status) {
  if (threadIdx.x == 0) {
    rank_out[blockIdx.x] = rank;
    resid_out[blockIdx.x] = rel;
    status_out[blockIdx.x] = status;
  }
}

__global__ __launch_bounds__(kThreads) void svd_qr_kernel(double* __restrict__ A,
                                                          const smx_svd_task* __restrict__ tasks, double rtol,
                                                          double* __restrict__ U, double* __restrict__ V,
                                                          double* __restrict__ work, int32_t* rank_out,
                                                          double* resid_out, int32_t* status_out) {
  __shared__ Shared sh;
  const smx_svd_task t = tasks[blockIdx.x];
  const int m = t.m, n = t.n;
  double* Ab = A + t.a_off;
  const double norm_sq = block_norm_sq(Ab, m, n, sh);
  if (norm_sq == 0.0 || t.mr <= 0) {
    finish_status(rank_out, resid_out, status_out, 0, 0.0, ST_ZERO);
    return;
  }
  const int mn = m < n ? m : n;
  const int kq = mn < kJMax ? mn : kJMax;
  double* ws = work + t.w_off;
  double* nrm = ws;             // m: ||X[j:, c]||^2 of the remaining rows of column c
  double* pivd = ws + m;        // m: original row of A in column slot c
  double* tau = ws + 2 * (int64_t)m;
  double* T = tau + kq;         // kq x m
  double* J = T + (int64_t)kq * m;
  for (int c = wave_id(); c < m; c += kWaves) {
    const double* ac = Ab + (int64_t)c * n;
    double s = 0.0;
    for (int l = lane_id(); l < n; l += 64) s = fma(ac[l], ac[l], s);
    s = smx::wave_total(s);
    if (lane_id() == 0) {
      nrm[c] = s;
      pivd[c] = (double)c;
    }
  }
  __syncthreads();
  const double floor_hi = (kFloorHi * rtol) * (kFloorHi * rtol) * norm_sq;
  const double floor_lo = (kFloorLo * rtol) * (kFloorLo * rtol) * norm_sq;
  int k = 0;
  bool reached = false;
  for (int j = 0; j <= kq; ++j) {
    double tot = 0.0, bv = -1.0;
    int bi = 0x7fffffff;
    for (int c = j + (int)threadIdx.x; c < m; c += kThreads) {
      const double x = nrm[c];
      tot += x;
      if (x > bv) {
        bv = x;
        bi = c;
      }
    }
    tot = block_sum(tot, sh);
    if (tot <= floor_hi || j == mn) {
      reached = true;
      break;
    }
    if (j == kq) {  // kq < mn steps done
      reached = tot <= floor_lo;
      break;
    }
    block_argmax(bv, bi, sh);
    const int p = bi;
    if (p != j) {
      double* aj = Ab + (int64_t)j * n;
      double* ap = Ab + (int64_t)p * n;
      for (int l = threadIdx.x; l < n; l += kThreads) {
        const double x = aj[l];
        aj[l] = ap[l];
        ap[l] = x;
      }
      if (threadIdx.x == 0) {
        const double x = nrm[j];
        nrm[j] = nrm[p];
        nrm[p] = x;
        const double y = pivd[j];
        pivd[j] = pivd[p];
        pivd[p] = y;
      }
    }
    __syncthreads();
    // reflector of column slot j over rows j .. n-1 (the LAPACK dlarfg form)
    double* xj = Ab + (int64_t)j * n;
    if (wave_id() == 0) {
      double xs = 0.0;
      for (int l = j + 1 + lane_id(); l < n; l += 64) xs = fma(xj[l], xj[l], xs);
      xs = smx::wave_total(xs);
      const double alpha = xj[j];
      double tj = 0.0, scale = 0.0, beta = alpha;
      if (xs != 0.0) {
        beta = -copysign(sqrt(fma(alpha, alpha, xs)), alpha);
        tj = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      if (xs != 0.0) {
        for (int l = j + 1 + lane_id(); l < n; l += 64) xj[l] *= scale;
      }
      if (lane_id() == 0) {
        xj[j] = beta;
        tau[j] = tj;
        sh.bcast[0] = tj;
      }
    }
    __syncthreads();
    const double tj = sh.bcast[0];
    for (int c = j + 1 + wave_id(); c < m; c += kWaves) {
      double* xc = Ab + (int64_t)c * n;
      double w = 0.0;
      for (int l = j + lane_id(); l < n; l += 64) w = fma(l == j ? 1.0 : xj[l], xc[l], w);
      w = smx::wave_total(w);
      const double tw = tj * w;
      double s = 0.0;
      for (int l = j + lane_id(); l < n; l += 64) {
        const double v = l == j ? 1.0 : xj[l];
        const double x = fma(-tw, v, xc[l]);
        xc[l] = x;
        if (l > j) s = fma(x, x, s);
      }
      s = smx::wave_total(s);
      if (lane_id() == 0) nrm[c] = s;
    }
    __syncthreads();
    k = j + 1;
  }
  if (!reached) {
    finish_status(rank_out, resid_out, status_out, 0, INFINITY, ST_FAILED);
    return;
  }
  if (k == 0) k = 1;  // unreachable for a nonzero block; keeps T nonempty
  // T: the k leading rows of the triangular factor, row-major k x m
  for (int64_t e = threadIdx.x; e < (int64_t)k * m; e += kThreads) {
    const int i = (int)(e / m), pos = (int)(e % m);
    T[e] = i <= pos ? Ab[(int64_t)pos * n + i] : 0.0;
  }
  set_identity(J, k);
  jacobi_rows(T, m, k, m, J, sh);
  row_norms_sorted(T, m, k, m, sh);
  if (threadIdx.x == 0) {
    const int usable = t.mr < mn ? t.mr : mn;
    int rank;
    double rel;
    const bool found = rank_rule(sh, k, usable, norm_sq, rtol, rank, rel);
    int status = ST_STORED;
    if (!found && usable > k) {
      // the floor is below rtol, so this is round-off at the threshold: the
      // k captured values are all there is
      rank = k;
    }
    if (rank > t.rcap) status = ST_OVER_CAP;
    sh.ibcast[0] = rank;
    sh.ibcast[1] = status;
    sh.bcast[1] = rel;
  }
  __syncthreads();
  const int rank = sh.ibcast[0];
  const int status = sh.ibcast[1];
  finish_status(rank_out, resid_out, status_out, rank, sh.bcast[1], status);
  if (status != ST_STORED || rank == 0) return;
  // U column r: row ord[r] of the rotated T, unpivoted (carries s_r)
  for (int64_t e = threadIdx.x; e < (int64_t)rank * m; e += kThreads) {
    const int r = (int)(e / m), pos = (int)(e % m);
    U[t.u_off + (int64_t)r * m + (int64_t)pivd[pos]] = T[(int64_t)sh.ord[r] * m + pos];
  }
  // V row r: Q_H [J(ord[r], :), 0]^T, the reflectors applied last to first;
  // lane l owns entries l, l + 64, ...
  for (int r = wave_id(); r < rank; r += kWaves) {
    double* vr = V + t.v_off + (int64_t)r * n;
    const double* jr = J + (int64_t)sh.ord[r] * k;
    for (int l = lane_id(); l < n; l += 64) vr[l] = l < k ? jr[l] : 0.0;
    for (int jj = k - 1; jj >= 0; --jj) {
      const double* xv = Ab + (int64_t)jj * n;
      double w = 0.0;
      for (int l = jj + lane_id(); l < n; l += 64) w = fma(l == jj ? 1.0 : xv[l], vr[l], w);
      w = smx::wave_total(w);
      const double tw = tau[jj] * w;
      for (int l = jj + lane_id(); l < n; l += 64) vr[l] = fma(-tw, l == jj ? 1.0 : xv[l], vr[l]);
    }
  }
}

__global__ __launch_bounds__(kThreads) void rsvd_round_kernel(const double* __restrict__ A,
                                                              const smx_svd_task* __restrict__ tasks,
                                                              const double* __restrict__ omega, double rtol,
                                                              int32_t power, double* __restrict__ U,
                                                              double* __restrict__ V, double* __restrict__ work,
                                                              int32_t* rank_out, double* resid_out,
                                                              int32_t* status_out) {
  __shared__ Shared sh;
  const smx_svd_task t = tasks[blockIdx.x];
  const int m = t.m, n = t.n, s = t.s;
  const double* Ab = A + t.a_off;
  const double norm_sq = block_norm_sq(Ab, m, n, sh);
  if (norm_sq == 0.0 || t.mr <= 0) {
    finish_status(rank_out, resid_out, status_out, 0, 0.0, ST_ZERO);
    return;
  }
  double* ws = work + t.w_off;
  double* Y = ws;                          // s x m: sketch rows, then Q^T
  double* Z = Y + (int64_t)s * m;          // s x n: Q^T A
  double* J = Z + (int64_t)s * n;          // s x s
  const double* Om = omega + t.om_off;     // n x s, row-major
  gemm_at(Om, 1, s, Ab, m, n, s, Y);       // Y^T = Omega^T A^T
  __syncthreads();
  for (int it = 0; it < power; ++it) {
    orth_rows(Y, s, m, sh);
    gemm_a(Y, Ab, m, n, s, Z);             // Q^T A
    __syncthreads();
    gemm_at(Z, n, 1, Ab, m, n, s, Y);      // (A A^T Q)^T
    __syncthreads();
  }
  orth_rows(Y, s, m, sh);
  gemm_a(Y, Ab, m, n, s, Z);               // B = Q^T A
  __syncthreads();
  set_identity(J, s);
  jacobi_rows(Z, n, s, n, J, sh);
  row_norms_sorted(Z, n, s, n, sh);
  if (threadIdx.x == 0) {
    const int usable = t.mr < s ? t.mr : s;
    int rank;
    double rel;
    const bool found = rank_rule(sh, s, usable, norm_sq, rtol, rank, rel);
    int status = ST_STORED;
    if (!found && s < t.mr) {
      status = ST_WIDER;
    } else if (rank > t.rcap) {
      status = ST_OVER_CAP;
    }
    sh.ibcast[0] = rank;
    sh.ibcast[1] = status;
    sh.bcast[1] = rel;
  }
  __syncthreads();
  const int rank = sh.ibcast[0];
  const int status = sh.ibcast[1];
  finish_status(rank_out, resid_out, status_out, status == ST_WIDER ? 0 : rank, sh.bcast[1], status);
  if (status != ST_STORED || rank == 0) return;
  // U column r = s_r Q J(ord[r], :)^T; V row r = row ord[r] of the rotated B / s_r
  for (int64_t e = threadIdx.x; e < (int64_t)rank * m; e += kThreads) {
    const int r = (int)(e / m), row = (int)(e % m);
    const int o = sh.ord[r];
    const double* jr = J + (int64_t)o * s;
    double acc = 0.0;
    for (int l = 0; l < s; ++l) acc = fma(jr[l], Y[(int64_t)l * m + row], acc);
    U[t.u_off + (int64_t)r * m + row] = acc * sh.sig[o];
  }
  for (int64_t e = threadIdx.x; e < (int64_t)rank * n; e += kThreads) {
    const int r = (int)(e / n), j = (int)(e % n);
    const int o = sh.ord[r];
    const double sg = sh.sig[o];
    V[t.v_off + (int64_t)r * n + j] = sg > 0.0 ? Z[(int64_t)o * n + j] / sg : 0.0;
  }
}

int check_tasks(const smx_svd_task* tasks, int64_t ntasks) {
  SMX_REQUIRE(ntasks >= 0 && ntasks <= 0x7fffffffLL, SMX_ERR_SHAPE, "task count %lld out of range",
              (long long)ntasks);
  SMX_REQUIRE(ntasks == 0 || tasks, SMX_ERR_VALUE, "null task table");
  return SMX_OK;
}

}  // namespace

extern "C" int64_t smx_svd_ws_doubles(int32_t m, int32_t n, int32_t s) {
  if (m <= 0 || n <= 0 || s < 0) return -1;
  if (s == 0) {
    const int64_t mn = m < n ? m : n;
    const int64_t kq = mn < kJMax ? mn : kJMax;
    return 2 * (int64_t)m + kq + kq * m + kq * kq;
  }
  if (s > kJMax) return -1;
  return (int64_t)s * m + (int64_t)s * n + (int64_t)s * s;
}

extern "C" int smx_svd_batched(double* A, const smx_svd_task* tasks, int64_t ntasks, double rtol, double* U,
                               double* V, double* work, int32_t* rank_out, double* resid_out,
                               int32_t* status_out, void* stream) {
  int rc = check_tasks(tasks, ntasks);
  if (rc) return rc;
  if (ntasks == 0) return SMX_OK;
  SMX_REQUIRE(A && U && V && work && rank_out && resid_out && status_out, SMX_ERR_VALUE,
              "null pointer argument");
  SMX_REQUIRE(rtol >= 0.0, SMX_ERR_VALUE, "rtol must be >= 0");
  hipLaunchKernelGGL(svd_qr_kernel, dim3((unsigned)ntasks), dim3(kThreads), 0, smx::as_stream(stream), A, tasks,
                     rtol, U, V, work, rank_out, resid_out, status_out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_rsvd_round(const double* A, const smx_svd_task* tasks, int64_t ntasks, const double* omega,
                              double rtol, int32_t power_iterations, double* U, double* V, double* work,
                              int32_t* rank_out, double* resid_out, int32_t* status_out, void* stream) {
  int rc = check_tasks(tasks, ntasks);
  if (rc) return rc;
  if (ntasks == 0) return SMX_OK;
  SMX_REQUIRE(A && omega && U && V && work && rank_out && resid_out && status_out, SMX_ERR_VALUE,
              "null pointer argument");
  SMX_REQUIRE(rtol >= 0.0 && power_iterations >= 0, SMX_ERR_VALUE, "rtol and power_iterations must be >= 0");
  hipLaunchKernelGGL(rsvd_round_kernel, dim3((unsigned)ntasks), dim3(kThreads), 0, smx::as_stream(stream), A,
                     tasks, omega, rtol, power_iterations, U, V, work, rank_out, resid_out, status_out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}
