// Conjugate gradients preconditioned by one smoothed-aggregation AMG V-cycle.
//
// The V-cycle (damped-Jacobi pre- and post-smoothing, one sweep each, zero
// initial guess on every level) is folded into three sparse operators per
// level, built once on the host:
//   S = 2w D^-1 - w^2 D^-1 A D^-1        (pre-smooth, residual, post-smooth)
//   Q = (I - w D^-1 A) P                 (prolongation, then post-smoothing)
//   R = Q^T                              (pre-smoothing, residual, restriction)
// so that M^-1 b = S b + Q M_c^-1 (R b), with M_c^-1 the V-cycle of the next
// level and, on the coarsest level, a dense pseudo-inverse (also stored as a
// sparse operator with full rows).  One level of the cycle is therefore one
// launch on the way down (b' = R b) and one on the way up (x = S b + Q x').
//
// CG in the Chronopoulos-Gear form (one reduction per iteration):
//   s = w + b s;  p = u + b p;  x += a p;  r -= a s;  u = M^-1 r;  w = A u
//   g = (r,u), d = (w,u), rr = (r,r);  b = g/g_old;  a = g/(d - b g/a_old)
// Launches of one iteration:
//   update   fine rows: s, p, x, r and partials of (rr, sum x);
//            extra workgroups: b_1 = R_0 r_new by linearity from carried
//            coarse vectors, b_1 = R_0 r and z_1 = R_0 s, as
//              z_1 <- R_0 w + b z_1,  b_1 <- b_1 - a z_1
//            (the same recurrence as s and r), so the fine level is gathered
//            once per R_0 entry (w) instead of three times (r, w, s);
//   down     b_{l+1} = R_l b_l                 (levels 1 .. L-2)
//   coarse   x_c = Z b_c
//   up       x_l = S_l b_l + Q_l x_{l+1}       (levels L-2 .. 1)
//   up0      u = S_0 r_new + Q_0 x_1
//   matvec   w = A u, partials of (w, u) and (r_new, u)
// Every workgroup of the next update re-reduces all partial sums in one fixed
// order (no atomics, no device-scope fences), so every workgroup derives the
// same scalars.  The update tests the residual of the previous iteration and
// raises a done flag; every later launch reads it first and returns at once.
//
// Level operators are sliced ELL with TPR lanes per row (TPR a power of two)
// and 256/TPR rows per slice, one slice per 256 threads: the slice offset is
// one scalar load, every row of a slice has the slice's width (padding slots
// hold value 0 and a valid column), so the column and value loads need no
// row pointer and every lane of a slice runs the same trip count.  Each lane
// sums every TPR-th slot of its row and the lanes of a row meet in a fixed
// butterfly, so results do not depend on scheduling.  Values are FP64 or FP32
// (the FP32 hierarchy halves the bytes of the preconditioner).
//
// Mixed precision (smx_pcg_amg_mixed): the fine-level iteration vectors
// (r, s, w, u, p) and the operator of w = A u are FP32; x, the coarse-level
// vectors, the scalars and the inner products stay FP64 (exact products,
// compensated sums).  Iterative refinement: after the FP32 solve the residual
// b - A x is recomputed from the FP64 operator and checked (two launches); a
// correction solve from the current x runs only when the check fails.
#include "smx_common.h"

namespace {

enum {
  SS_DONE = 0,
  SS_ITERS = 1,
  SS_RR = 2,
  SS_TOL2 = 3,
  SS_BB = 4,
  SS_FAIL = 5,  // 1 breakdown, 2 max_it reached
  SS_SUMX = 6,
  SS_ITERS0 = 7,  // iterations of the inner solves before the last (mixed precision)
  SS_SIZE = 8,
  ST_OFF = 8,   // [gamma, alpha, 1/gamma, 1/alpha] x 2 parities
  ST_SIZE = 8,
  HDR = 32
};
static_assert(SS_DONE == SMX_PCG_DONE && SS_ITERS == SMX_PCG_ITERS && SS_FAIL == SMX_PCG_FAIL &&
                  SS_SUMX == SMX_PCG_SUMX,
              "status block layout of the C ABI");

// Fine-level iteration vectors are FP64 or FP32 (a kernel template
// parameter); x and the coarse vectors are FP64.
struct AmgPtrs {
  double* ss;
  double* st[2];
  double* pbb;      // Gf partials of ||b||^2 (written once, by the init launch)
  double* prr[2];   // 2 x Gf partials (rr, sum x), by iteration parity
  double* pg;       // Gf partials of (r, u)
  double* pd;       // Gf partials of (w, u)
  void* r[2];
  void* s[2];
  void* w;
  void* u;
  void* p;
  double* z1;       // R_0 s (coarse, carried)
  int Gf;           // fine workgroups (one row per thread)
};

template <typename V>
__host__ __device__ __forceinline__ V* vec(void* p) {
  return static_cast<V*>(p);
}

// the vector types of a launch are given explicitly, not deduced
template <typename T>
struct Ident {
  using type = T;
};

struct SellOp {
  const int32_t* off;  // slice offsets (nslices + 1)
  const int32_t* col;  // NULL: dense rows, the column of slot j is j
  const void* val;
  int64_t rows;
};

constexpr int kSliceThreads = 256;  // TPR * rows per slice

// Workgroup b of G is dispatched to XCD b % 8; renumber so that each XCD
// handles one contiguous range of blocks (bijective for every G).
__device__ __forceinline__ int xcd_map(int b, int G) {
  const int x = b & 7, k = b >> 3;
  const int q = G >> 3, r = G & 7;
  return x * q + (x < r ? x : r) + k;
}

// COMP: the sums across waves (and, in the callers, across workgroups' partials)
// are compensated (TwoSum); v[q] + c[q] is the per-thread input.  The wave
// totals stay a plain DPP butterfly: it sits on every iteration's dependent
// chain, and its inputs are FP64 sums of products that are exact in FP64.
template <int BS, int NQ, bool ALL, bool COMP = false>
__device__ __forceinline__ void block_total(double (&v)[NQ], double* scr, const double* c = nullptr) {
  constexpr int nw = BS / 64;
  const int wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = smx::wave_total(COMP && c ? v[q] + c[q] : v[q]);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) scr[wid * NQ + q] = v[q];
  }
  smx::lds_barrier();
  if (ALL || threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double s = 0.0, e = 0.0;
#pragma unroll
      for (int w = 0; w < nw; ++w) {
        if (COMP) {
          smx::comp_add(s, e, scr[w * NQ + q]);
        } else {
          s += scr[w * NQ + q];
        }
      }
      v[q] = COMP ? s + e : s;
    }
  }
}

template <int BS>
__device__ __forceinline__ double sum_parts(const double* __restrict__ part, int G) {
  double t = 0.0;
  for (int b = threadIdx.x; b < G; b += BS) t += part[b];
  return t;
}

template <typename VT>
__device__ __forceinline__ double ldv(const void* v, int64_t k) {
  return (double)static_cast<const VT*>(v)[k];
}

// Row i (local to a slice at `off` of width W) of a sliced-ELL operator times
// v over TPR lanes; every lane of the row ends with the total (fixed
// butterfly order).  Lane l takes slots l, l+TPR, ... in that order, four at a
// time: the four column and value loads, then the four gathers, are issued
// before the first multiply (masked slots re-read a valid slot and add an
// exact zero).  Slots are column-major in the slice for TPR == 1 (lanes of a
// wavefront read consecutive rows) and row-major otherwise (lanes of a row
// read consecutive slots).
template <int TPR, typename VT, typename XT>
__device__ __forceinline__ double sell_row(const SellOp& M, const XT* __restrict__ v, int32_t off,
                                           int32_t W, int i, int lane) {
  constexpr int U = 4;
  constexpr int RPS = kSliceThreads / TPR;
  double acc = 0.0;
  for (int32_t jb = lane; jb < W; jb += U * TPR) {
    double vv[U], g[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t j = jb + u * TPR;
      const bool ok = j < W;
      const int32_t jj = ok ? j : jb;
      const int32_t pos = TPR == 1 ? off + jj * RPS + i : off + i * W + jj;
      const int32_t c = M.col ? M.col[pos] : jj;
      vv[u] = ok ? ldv<VT>(M.val, pos) : 0.0;
      g[u] = (double)v[c];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc = fma(vv[u], g[u], acc);
  }
#pragma unroll
  for (int o = TPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  return acc;
}

// r_{i+1} = r_i - a (w_i + b s_{i-1}): one formula for the fine update and the
// fused restriction, so both see the same residual bit for bit.
__device__ __forceinline__ double next_s(double w, double b, double s) { return fma(b, s, w); }
__device__ __forceinline__ double next_r(double r, double a, double sn) { return fma(-a, sn, r); }

// r0 = b - A x0 (mode 0) or the lifted residual (mode 1); s = p = 0.
// Partials: ||b||^2 (pbb) and (||r0||^2, sum x0) (prr[0]).  Block 0 resets the
// status record.
template <int BS, int WMAX, typename V>
__global__ __launch_bounds__(BS) void amg_init_kernel(
    int64_t n, int64_t nd, int32_t mode, int32_t W, int64_t ld, const int32_t* __restrict__ cols,
    const double* __restrict__ vals, const double* __restrict__ b,
    const double* __restrict__ b_part, int64_t nb_part, int32_t project,
    const int32_t* __restrict__ bpos, const double* __restrict__ gb, double* __restrict__ x,
    int32_t warm, const int32_t* __restrict__ perm, AmgPtrs P, int64_t n1, int32_t phase) {
  __shared__ double scr[BS / 64];
  __shared__ double scr_pub[BS / 64 * 3];
  const int blk = xcd_map(blockIdx.x, P.Gf);
  const int64_t row = (int64_t)blk * BS + threadIdx.x;
  if (row < n1) P.z1[row] = 0.0;  // R_0 s_{-1} = 0
  double shift = 0.0;
  if (mode == 0 && project && b_part) {
    double t[1] = {0.0};
    for (int64_t k = threadIdx.x; k < nb_part; k += BS) t[0] += b_part[k];
    block_total<BS, 1, true>(t, scr);
    shift = t[0] / (double)n;
  }
  double v[3] = {0.0, 0.0, 0.0};
  if (row < n) {
    double r0, bk, x0 = 0.0;
    if (mode == 0) {
      bk = b[perm ? perm[row] : row] - shift;
      double ax = 0.0;
      if (warm) {
        x0 = x[row];
#pragma unroll
        for (int j = 0; j < WMAX; ++j) {
          if (j < W) ax = fma(vals[j * ld + row], x[cols[j * ld + row]], ax);
        }
      }
      r0 = bk - ax;
    } else {
      const int32_t bk_pos = bpos[row];
      if (bk_pos >= 0) {
        bk = gb[bk_pos];
        x0 = bk;
        r0 = 0.0;
      } else {
        double sx = 0.0, sg = 0.0;
#pragma unroll
        for (int j = 0; j < WMAX; ++j) {
          if (j < W) {
            const int32_t c = cols[j * ld + row];
            const double a = vals[j * ld + row];
            const int32_t cp = bpos[c];
            if (cp >= 0) {
              const double gv = gb[cp];
              sx = fma(a, gv, sx);
              sg = fma(a, gv, sg);
            } else if (warm) {
              sx = fma(a, x[c], sx);
            }
          }
        }
        if (warm) x0 = x[row];
        r0 = -sx;
        bk = -sg;
      }
    }
    // Other workgroups may read x[row] as a neighbour of a warm start, so x is
    // written only where no workgroup reads it: cold starts, and the boundary
    // rows of mode 1 (their neighbours read gb instead).
    if (!warm && (mode == 0 || bpos[row] < 0)) x[row] = 0.0;
    if (mode == 1 && bpos[row] >= 0) x[row] = x0;
    const V r0v = (V)r0;
    vec<V>(P.r[0])[row] = r0v;
    vec<V>(P.s[0])[row] = (V)0.0;
    vec<V>(P.p)[row] = (V)0.0;
    v[0] = bk * bk;
    v[1] = (double)r0v * (double)r0v;
    v[2] = x0;
  }
  // trailing identity rows n .. n+nd-1: x is their data (b, or gb at bpos),
  // their residual is 0 and they count in ||b|| as rows of the full system
  for (int64_t t = n + row; t < n + nd; t += (int64_t)gridDim.x * BS) {
    const double bt = mode == 0 ? b[perm ? perm[t] : t] : gb[bpos[t]];
    x[t] = bt;
    v[0] = fma(bt, bt, v[0]);
  }
  block_total<BS, 3, false>(v, scr_pub);
  if (threadIdx.x == 0) {
    P.pbb[blk] = v[0];
    P.prr[0][blk] = v[1];
    P.prr[0][P.Gf + blk] = v[2];
  }
  // phase > 0 (a refinement check): the status is the check kernel's to write
  if (phase == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    P.ss[SS_ITERS0] = 0.0;
    P.ss[SS_DONE] = 0.0;
    P.ss[SS_FAIL] = 0.0;
    P.ss[SS_ITERS] = 0.0;
  }
}

// Refinement check, one workgroup: the FP64 residual ||b - A x|| left by
// amg_init_kernel against the tolerance.  Met: the solve is done.  Not met:
// the iterations so far move to SS_ITERS0 and the status is reset for a
// correction solve from the current x.
template <int BS>
__global__ __launch_bounds__(BS) void amg_check_kernel(AmgPtrs P, double rtol, double atol) {
  __shared__ double scr[BS / 64 * 3];
  const int Gf = P.Gf;
  double t[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < Gf; b += BS) {
    smx::comp_add(t[0], c[0], P.pbb[b]);
    smx::comp_add(t[1], c[1], P.prr[0][b]);
    smx::comp_add(t[2], c[2], P.prr[0][Gf + b]);
  }
  block_total<BS, 3, false, true>(t, scr, c);
  if (threadIdx.x == 0) {
    const double tol2 = fmax(rtol * rtol * t[0], atol * atol);
    P.ss[SS_BB] = t[0];
    P.ss[SS_TOL2] = tol2;
    P.ss[SS_RR] = t[1];
    P.ss[SS_SUMX] = t[2];
    P.ss[SS_FAIL] = 0.0;
    if (t[1] <= tol2) {
      P.ss[SS_DONE] = 1.0;
    } else {
      P.ss[SS_ITERS0] = P.ss[SS_ITERS0] + P.ss[SS_ITERS];
      P.ss[SS_ITERS] = 0.0;
      P.ss[SS_DONE] = 0.0;
    }
  }
}

// One level launch: y = M1 v1 (+ M2 v2), one slice of both operators per
// workgroup (M1 and M2 have the same rows and slicing).
template <int TPR, typename VT, typename X1, typename X2, typename Y>
__global__ __launch_bounds__(kSliceThreads) void amg_level_kernel(SellOp M1,
                                                                  const X1* __restrict__ v1,
                                                                  SellOp M2,
                                                                  const X2* __restrict__ v2,
                                                                  Y* __restrict__ y,
                                                                  const double* __restrict__ ss) {
  constexpr int RPS = kSliceThreads / TPR;
  // the done flag (written by an earlier launch) is tested once the slice
  // offsets are in flight
  const bool was_done = ss[SS_DONE] != 0.0;
  const int sl = xcd_map(blockIdx.x, gridDim.x);
  const int32_t a0 = M1.off[sl], a1 = M1.off[sl + 1];
  int32_t b0 = 0, b1 = 0;
  if (M2.off) {
    b0 = M2.off[sl];
    b1 = M2.off[sl + 1];
  }
  if (was_done) return;  // uniform
  const int i = threadIdx.x / TPR;
  const int lane = threadIdx.x % TPR;
  const int64_t row = (int64_t)sl * RPS + i;
  double acc = sell_row<TPR, VT>(M1, v1, a0, (a1 - a0) / RPS, i, lane);
  if (M2.off) acc += sell_row<TPR, VT>(M2, v2, b0, (b1 - b0) / RPS, i, lane);
  if (lane == 0 && row < M1.rows) y[row] = (Y)acc;
}

// Start of iteration `it`: scalars from the partial sums, convergence test,
// then (blocks < Gf) the fine update and (blocks >= Gf) b_1 = R_0 r_new, BS/256
// slices of R_0 per workgroup.  Everything that does not need the scalars
// (the fine vectors, and all of R_0 w) is loaded or computed before the
// reduction, so the reduction is the only wait on the partial sums.
template <int BS, int TPR, typename VT, typename V, bool COMP>
__global__ __launch_bounds__(BS) void amg_update_kernel(int64_t n, AmgPtrs P, double* __restrict__ x,
                                                        SellOp R0, double* __restrict__ b1,
                                                        int32_t it, int32_t max_it, double rtol,
                                                        double atol) {
  __shared__ double scr[BS / 64 * 5];
  __shared__ double scr_pub[BS / 64 * 2];
  const bool was_done = P.ss[SS_DONE] != 0.0;
  const int q = it & 1;
  const int Gf = P.Gf;
  const bool fine = (int)blockIdx.x < Gf;
  // partial sums first: the reduction waits only on these
  const int b0 = (int)threadIdx.x < Gf ? (int)threadIdx.x : 0;
  double t[5], tc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  t[0] = P.pg[b0];
  t[1] = P.pd[b0];
  t[2] = P.prr[q][b0];
  t[3] = P.prr[q][Gf + b0];
  t[4] = P.pbb[b0];
  const V* __restrict__ rq = vec<V>(P.r[q]);
  const V* __restrict__ sq = vec<V>(P.s[q]);
  const V* __restrict__ w = vec<V>(P.w);
  // fine role: the row's vectors; restriction role: (R_0 w)_j and b_1, z_1
  const int blk = fine ? xcd_map(blockIdx.x, Gf) : 0;
  const int64_t row = fine ? (int64_t)blk * BS + threadIdx.x
                           : (int64_t)(blockIdx.x - Gf) * (BS / TPR) + threadIdx.x / TPR;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0, f4 = 0.0, f5 = 0.0;
  if (fine) {
    const int64_t rc = row < n ? row : n - 1;
    f0 = (double)w[rc];
    f1 = (double)sq[rc];
    f2 = (double)vec<V>(P.p)[rc];
    f3 = (double)vec<V>(P.u)[rc];
    f4 = x[rc];
    f5 = (double)rq[rc];
  } else {
    constexpr int RPS = kSliceThreads / TPR;
    const int sl = (int)(row / RPS);
    const int i = (int)(row % RPS);
    const int nsl = (int)((R0.rows + RPS - 1) / RPS);
    if (sl < nsl) {  // uniform per slice (a workgroup may run past the last one)
      const int32_t o0 = R0.off[sl], o1 = R0.off[sl + 1];
      // a launch past convergence (speculative) skips the product: the done
      // flag was loaded with the slice offsets, so real iterations wait no longer
      if (!was_done) f0 = sell_row<TPR, VT>(R0, w, o0, (o1 - o0) / RPS, i, (int)threadIdx.x % TPR);
    }
    const int64_t rc = row < R0.rows ? row : R0.rows - 1;
    f1 = P.z1[rc];
    f2 = b1[rc];
  }
  if ((int)threadIdx.x >= Gf) {
#pragma unroll
    for (int k = 0; k < 5; ++k) t[k] = 0.0;
  }
  for (int b = threadIdx.x + BS; b < Gf; b += BS) {  // only when Gf > BS
    const double e[5] = {P.pg[b], P.pd[b], P.prr[q][b], P.prr[q][Gf + b], P.pbb[b]};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (COMP) {
        smx::comp_add(t[k], tc[k], e[k]);
      } else {
        t[k] += e[k];
      }
    }
  }
  block_total<BS, 5, true, COMP>(t, scr, tc);
  const double gam = t[0], del = t[1], rr = t[2];
  const double tol2 = fmax(rtol * rtol * t[4], atol * atol);
  double alpha, beta;
  if (it == 0) {
    beta = 0.0;
    alpha = gam / del;
  } else {
    const double inv_g_old = P.st[q][2], inv_a_old = P.st[q][3];
    beta = gam * inv_g_old;
    alpha = gam / (del - beta * gam * inv_a_old);
  }
  const bool conv = rr <= tol2;
  const bool fail = !conv && !(gam > 0.0 && isfinite(alpha));
  if (was_done) return;  // uniform
  const bool stop = conv || fail || it >= max_it;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    P.ss[SS_ITERS] = (double)it;
    P.ss[SS_RR] = rr;
    P.ss[SS_SUMX] = t[3];
    P.ss[SS_BB] = t[4];
    P.ss[SS_TOL2] = tol2;
    if (stop) {
      P.ss[SS_DONE] = 1.0;
      P.ss[SS_FAIL] = fail ? 1.0 : (conv ? 0.0 : 2.0);
    } else {
      P.st[1 - q][0] = gam;
      P.st[1 - q][1] = alpha;
      P.st[1 - q][2] = 1.0 / gam;
      P.st[1 - q][3] = 1.0 / alpha;
    }
  }
  if (stop) return;
  if (fine) {
    double v[2] = {0.0, 0.0};
    if (row < n) {
      const V s_n = (V)next_s(f0, beta, f1);
      const V p_n = (V)fma(beta, f2, f3);
      const double x_n = fma(alpha, (double)p_n, f4);
      const V r_n = (V)next_r(f5, alpha, (double)s_n);
      __builtin_nontemporal_store(s_n, &vec<V>(P.s[1 - q])[row]);
      __builtin_nontemporal_store(r_n, &vec<V>(P.r[1 - q])[row]);
      __builtin_nontemporal_store(p_n, &vec<V>(P.p)[row]);
      __builtin_nontemporal_store(x_n, &x[row]);
      v[0] = (double)r_n * (double)r_n;
      v[1] = x_n;
    }
    block_total<BS, 2, false, COMP>(v, scr_pub);
    if (threadIdx.x == 0) {
      P.prr[1 - q][blk] = v[0];
      P.prr[1 - q][Gf + blk] = v[1];
    }
  } else if (row < R0.rows && threadIdx.x % TPR == 0) {
    const double z_n = next_s(f0, beta, f1);
    b1[row] = next_r(f2, alpha, z_n);
    P.z1[row] = z_n;
  }
}

// w = A u (plain operator, ELL), partials of (w, u) and (r, u), r the
// residual the preconditioner was applied to.
template <int BS, int WMAX, bool EXACT, typename V, typename AV, bool COMP>
__global__ __launch_bounds__(BS) void amg_matvec_kernel(int64_t n, int32_t W, int64_t ld,
                                                        const int32_t* __restrict__ cols,
                                                        const AV* __restrict__ vals, AmgPtrs P,
                                                        const V* __restrict__ r) {
  __shared__ double scr[BS / 64 * 2];
  const bool was_done = P.ss[SS_DONE] != 0.0;  // tested once the matrix loads are in flight
  const int blk = xcd_map(blockIdx.x, P.Gf);
  const int64_t row = (int64_t)blk * BS + threadIdx.x;
  const int64_t rowc = row < n ? row : n - 1;
  const int Wk = EXACT ? WMAX : W;
  int32_t c[WMAX];
  double av[WMAX];
#pragma unroll
  for (int j = 0; j < WMAX; ++j) {
    c[j] = j < Wk ? cols[j * ld + rowc] : (int32_t)rowc;
    av[j] = j < Wk ? (double)vals[j * ld + rowc] : 0.0;
  }
  if (was_done) return;  // uniform
  double t[2] = {0.0, 0.0};
  if (row < n) {
    const V* __restrict__ u = vec<V>(P.u);
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < WMAX; ++j) acc = fma(av[j], (double)u[c[j]], acc);
    const V w_n = (V)acc;
    __builtin_nontemporal_store(w_n, &vec<V>(P.w)[row]);
    const double uk = (double)u[row];
    t[0] = (double)w_n * uk;
    t[1] = (double)r[row] * uk;
  }
  block_total<BS, 2, false, COMP>(t, scr);
  if (threadIdx.x == 0) {
    P.pd[blk] = t[0];
    P.pg[blk] = t[1];
  }
}

__global__ void amg_gather_kernel(int64_t n, const int32_t* __restrict__ perm,
                                  const double* __restrict__ x, double* __restrict__ xp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xp[i] = x[perm[i]];
}

__global__ void amg_scatter_kernel(int64_t n, const int32_t* __restrict__ perm,
                                   const double* __restrict__ xp, double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[perm[i]] = xp[i];
}

int fine_bs(int64_t n) { return n <= 65536 ? 256 : 512; }

SellOp to_op(const smx_amg_op& o) {
  SellOp c;
  c.off = o.slice_off;
  c.col = o.col;
  c.val = o.val;
  c.rows = o.rows;
  return c;
}

int64_t slices(int64_t rows, int tpr) {
  const int64_t rps = kSliceThreads / tpr;
  return (rows + rps - 1) / rps;
}

template <typename VT, typename X1 = double, typename X2 = double, typename Y = double>
int launch_level(const smx_amg_op& m1, const typename Ident<X1>::type* v1, const smx_amg_op* m2,
                 const typename Ident<X2>::type* v2, typename Ident<Y>::type* y, const double* ss,
                 int tpr, hipStream_t st) {
  const SellOp a = to_op(m1);
  SellOp b = {nullptr, nullptr, nullptr, 0};
  if (m2 && m2->rows > 0) b = to_op(*m2);
  if (a.rows <= 0) return SMX_OK;
  const unsigned g = (unsigned)slices(a.rows, tpr);
#define SMX_AMG_LVL(T)                                                                          \
  hipLaunchKernelGGL((amg_level_kernel<T, VT, X1, X2, Y>), dim3(g), dim3(kSliceThreads), 0, st, \
                     a, v1, b, v2, y, ss)
  switch (tpr) {
    case 1: SMX_AMG_LVL(1); break;
    case 2: SMX_AMG_LVL(2); break;
    case 4: SMX_AMG_LVL(4); break;
    case 8: SMX_AMG_LVL(8); break;
    case 16: SMX_AMG_LVL(16); break;
    case 32: SMX_AMG_LVL(32); break;
    default: SMX_AMG_LVL(64); break;
  }
#undef SMX_AMG_LVL
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

struct Layout {
  int64_t Gf, nvec;
};

// Gf from the n iterated rows; vectors (x in solver order among them) hold
// ntot = n + nd rows, so smx_pcg_amg_ws_size(ntot) covers the layout
Layout layout(int64_t n, int64_t ntot) {
  Layout L;
  L.Gf = (n + fine_bs(n) - 1) / fine_bs(n);
  L.nvec = (ntot + 63) / 64 * 64;
  return L;
}

bool valid_tpr(int t) { return t == 1 || t == 2 || t == 4 || t == 8 || t == 16 || t == 32 || t == 64; }

// The V-cycle below the fine level and the up0 launch (u = M^-1 r_fine).
// On entry levels[1].b holds R_0 r_fine (or, with one level, nothing).
template <typename VT, typename V>
int vcycle_tail(const smx_amg_level* lv, int nlev, const V* r_fine, AmgPtrs& P,
                hipStream_t st) {
  int rc;
  for (int l = 1; l + 1 < nlev; ++l) {  // down: b_{l+1} = R_l b_l
    rc = launch_level<VT>(lv[l].R, lv[l].b, nullptr, nullptr, lv[l + 1].b, P.ss, lv[l].r_tpr, st);
    if (rc) return rc;
  }
  if (nlev > 1) {  // coarse: x_c = Z b_c
    const smx_amg_level& c = lv[nlev - 1];
    rc = launch_level<VT>(c.S, c.b, nullptr, nullptr, c.x, P.ss, c.s_tpr, st);
    if (rc) return rc;
  }
  for (int l = nlev - 2; l >= 1; --l) {  // up: x_l = S_l b_l + Q_l x_{l+1}
    rc = launch_level<VT>(lv[l].S, lv[l].b, &lv[l].Q, lv[l + 1].x, lv[l].x, P.ss, lv[l].s_tpr,
                          st);
    if (rc) return rc;
  }
  const smx_amg_op* q0 = nlev > 1 ? &lv[0].Q : nullptr;
  const double* x1 = nlev > 1 ? lv[1].x : nullptr;
  return launch_level<VT, V, double, V>(lv[0].S, r_fine, q0, x1, vec<V>(P.u), P.ss, lv[0].s_tpr, st);
}

// The solve: AMG-CG from x0 (warm) or 0 to rtol_inner (rtol without
// refinement; rtol_inner may be below rtol), then up to `refinements` times the FP64 residual check
// (amg_init_kernel + amg_check_kernel) and, if it fails, a correction solve
// from the current x.  A speculative launch enqueues the first solve and one
// check: a failed check leaves DONE = 0 and the caller solves again waiting.
template <int BS, int WMAX, bool EXACT, typename VT, typename V, typename AV, bool COMP>
int run_amg(int64_t n, int64_t nd, int32_t W, int64_t ld, const int32_t* cols, const double* vals,
            const AV* ivals, int32_t mode, const double* b, const double* b_part, int64_t nb_part,
            int32_t project, int32_t raw_W, int64_t raw_ld, const int32_t* raw_cols,
            const double* raw_vals, const int32_t* bpos, const double* gb, double* x_user,
            int32_t warm, const int32_t* perm, double* x_perm, const smx_amg_level* lv, int nlev,
            double rtol, double atol, int32_t max_it, int32_t check_every, int32_t expected_its,
            int32_t refinements, double rtol_inner, AmgPtrs P, hipStream_t st, double* hst) {
  const int Gf = P.Gf;
  const unsigned gs = (unsigned)((n + nd + 255) / 256);
  double* x = perm ? x_perm : x_user;
  const int64_t n1 = nlev > 1 ? lv[0].R.rows : 0;
  if (perm && warm) {
    hipLaunchKernelGGL(amg_gather_kernel, dim3(gs), dim3(256), 0, st, n + nd, perm, x_user, x);
    SMX_LAUNCH_CHECK();
  }
  const int rt = nlev > 1 ? lv[0].r_tpr : 1;
  const int64_t Gr = nlev > 1 ? (n1 + (BS / rt) - 1) / (BS / rt) : 0;
  SellOp R0 = {nullptr, nullptr, nullptr, 0};
  if (nlev > 1) R0 = to_op(lv[0].R);
  double* b1 = nlev > 1 ? lv[1].b : nullptr;
  // check_every < 0: speculative launch of expected + 1 - check_every
  // iterations with no wait (iterations past convergence return at once on
  // the done flag); the caller reads the status block later
  const bool spec = check_every < 0 && expected_its > 0;
  const int32_t every = check_every < 1 ? 8 : check_every;
  auto init = [&](int32_t w, int32_t phase) -> int {
    if (mode == 0) {
      hipLaunchKernelGGL((amg_init_kernel<BS, WMAX, V>), dim3(Gf), dim3(BS), 0, st, n, nd, mode, W, ld,
                         cols, vals, b, b_part, nb_part, project, bpos, gb, x, w, perm, P, n1, phase);
    } else {
      hipLaunchKernelGGL((amg_init_kernel<BS, WMAX, V>), dim3(Gf), dim3(BS), 0, st, n, nd, mode, raw_W,
                         raw_ld, raw_cols, raw_vals, b, b_part, nb_part, project, bpos, gb, x, w, perm,
                         P, n1, phase);
    }
    SMX_LAUNCH_CHECK();
    return SMX_OK;
  };
  // one AMG-CG solve on the residual in P.r[0] (left by amg_init_kernel)
  auto iterate = [&](double tol, int32_t expect) -> int {
    int rc;
    // u0 = M^-1 r0, w0 = A u0
    const V* r0 = vec<V>(P.r[0]);
    if (nlev > 1) {
      rc = launch_level<VT, V, double, double>(lv[0].R, r0, nullptr, nullptr, lv[1].b, P.ss,
                                               lv[0].r_tpr, st);
      if (rc) return rc;
    }
    rc = vcycle_tail<VT, V>(lv, nlev, r0, P, st);
    if (rc) return rc;
    hipLaunchKernelGGL((amg_matvec_kernel<BS, WMAX, EXACT, V, AV, COMP>), dim3(Gf), dim3(BS), 0, st,
                       n, W, ld, cols, ivals, P, r0);
    SMX_LAUNCH_CHECK();
    int32_t batch = spec ? expect + 1 - check_every : (expect > 0 ? expect + 1 : every);
    int32_t launched = 0;
    bool done = false;
    while (!done && launched <= max_it) {
      const int32_t nb = batch < (max_it + 1 - launched) ? batch : (max_it + 1 - launched);
      for (int32_t k = 0; k < nb; ++k) {
        const int32_t it = launched + k;
        const unsigned grid = (unsigned)(Gf + Gr);
#define SMX_AMG_UPD(T)                                                                               \
  hipLaunchKernelGGL((amg_update_kernel<BS, T, VT, V, COMP>), dim3(grid), dim3(BS), 0, st, n, P, x, \
                     R0, b1, it, max_it, tol, atol)
        switch (rt) {
          case 1: SMX_AMG_UPD(1); break;
          case 2: SMX_AMG_UPD(2); break;
          case 4: SMX_AMG_UPD(4); break;
          case 8: SMX_AMG_UPD(8); break;
          case 16: SMX_AMG_UPD(16); break;
          case 32: SMX_AMG_UPD(32); break;
          default: SMX_AMG_UPD(64); break;
        }
#undef SMX_AMG_UPD
        SMX_LAUNCH_CHECK();
        // the last update of a speculative batch can only find convergence (if
        // it does not, the caller solves again): the V-cycle and A u after it
        // would never be used
        if (spec && k == nb - 1) continue;
        const V* rn = vec<V>(P.r[1 - (it & 1)]);
        rc = vcycle_tail<VT, V>(lv, nlev, rn, P, st);
        if (rc) return rc;
        hipLaunchKernelGGL((amg_matvec_kernel<BS, WMAX, EXACT, V, AV, COMP>), dim3(Gf), dim3(BS), 0,
                           st, n, W, ld, cols, ivals, P, rn);
        SMX_LAUNCH_CHECK();
      }
      launched += nb;
      if (spec) break;
      SMX_HIP_CHECK(hipMemcpyAsync(hst, P.ss, sizeof(double) * SS_SIZE, hipMemcpyDeviceToHost, st));
      SMX_HIP_CHECK(hipStreamSynchronize(st));
      done = hst[SS_DONE] != 0.0;
      batch = every;
    }
    return SMX_OK;
  };
  int rc = init(warm, 0);
  if (rc) return rc;
  rc = iterate(refinements > 0 ? rtol_inner : rtol, expected_its);
  if (rc) return rc;
  for (int32_t k = 1; k <= refinements; ++k) {
    rc = init(1, k);
    if (rc) return rc;
    hipLaunchKernelGGL((amg_check_kernel<1024>), dim3(1), dim3(1024), 0, st, P, rtol, atol);
    SMX_LAUNCH_CHECK();
    if (spec) break;
    SMX_HIP_CHECK(hipMemcpyAsync(hst, P.ss, sizeof(double) * SS_SIZE, hipMemcpyDeviceToHost, st));
    SMX_HIP_CHECK(hipStreamSynchronize(st));
    if (hst[SS_DONE] != 0.0) break;
    // correction k stops at rtol_inner^(k+1) of ||b||, the last at rtol
    rc = iterate(k == refinements ? rtol : fmax(rtol, pow(rtol_inner, (double)(k + 1))), 0);
    if (rc) return rc;
  }
  if (perm) {
    hipLaunchKernelGGL(amg_scatter_kernel, dim3(gs), dim3(256), 0, st, n + nd, perm, x, x_user);
    SMX_LAUNCH_CHECK();
  }
  if (!spec) {
    SMX_HIP_CHECK(hipMemcpyAsync(hst, P.ss, sizeof(double) * SS_SIZE, hipMemcpyDeviceToHost, st));
    SMX_HIP_CHECK(hipStreamSynchronize(st));
  }
  return SMX_OK;
}

__global__ void amg_clear_status_kernel(double* ss) {
  if (threadIdx.x < SS_SIZE) ss[threadIdx.x] = 0.0;
}

int check_levels(const smx_amg_level* levels, int32_t nlev, int64_t n) {
  SMX_REQUIRE(levels && nlev >= 1 && nlev <= 16, SMX_ERR_VALUE, "1 to 16 AMG levels required");
  // level l has n_l unknowns; S_l is n_l x n_l, Q_l n_l x n_{l+1}, R_l n_{l+1} x n_l;
  // the last level holds only S (the coarse inverse)
  SMX_REQUIRE(levels[0].n == n, SMX_ERR_SHAPE, "level 0 has %lld unknowns, the operator %lld",
              (long long)levels[0].n, (long long)n);
  for (int l = 0; l < nlev; ++l) {
    const smx_amg_level& L = levels[l];
    SMX_REQUIRE(L.n >= 1 && L.S.rows == L.n && L.S.slice_off && L.S.val, SMX_ERR_SHAPE,
                "level %d: S must have n rows", l);
    SMX_REQUIRE(valid_tpr(L.s_tpr), SMX_ERR_VALUE,
                "level %d: lanes per row must be 1..64, a power of 2", l);
    if (l > 0) SMX_REQUIRE(L.b && L.x, SMX_ERR_VALUE, "level %d: b and x vectors required", l);
    if (l + 1 < nlev) {
      SMX_REQUIRE(L.Q.rows == L.n && L.R.rows == levels[l + 1].n && L.Q.slice_off && L.Q.col &&
                      L.Q.val && L.R.slice_off && L.R.col && L.R.val,
                  SMX_ERR_SHAPE, "level %d: Q must have n_l rows, R n_{l+1} rows", l);
      SMX_REQUIRE(valid_tpr(L.r_tpr), SMX_ERR_VALUE,
                  "level %d: lanes per row must be 1..64, a power of 2", l);
    }
  }
  return SMX_OK;
}

template <typename VT>
int apply_once(const smx_amg_level* lv, int nlev, const double* r, double* u, double* work,
               hipStream_t st) {
  AmgPtrs P = {};
  P.ss = work;
  P.u = u;
  hipLaunchKernelGGL(amg_clear_status_kernel, dim3(1), dim3(64), 0, st, P.ss);
  SMX_LAUNCH_CHECK();
  if (nlev > 1) {
    const int rc = launch_level<VT>(lv[0].R, r, nullptr, nullptr, lv[1].b, P.ss, lv[0].r_tpr, st);
    if (rc) return rc;
  }
  return vcycle_tail<VT, double>(lv, nlev, r, P, st);
}

}  // namespace

extern "C" int smx_amg_apply(const smx_amg_level* levels, int32_t nlev, int32_t val_dtype,
                             const double* r, double* u, double* work, void* stream) {
  SMX_REQUIRE(val_dtype == SMX_F64 || val_dtype == SMX_F32, SMX_ERR_TYPE,
              "AMG operator values must be FP64 or FP32");
  SMX_REQUIRE(levels, SMX_ERR_VALUE, "null levels");
  const int rc = check_levels(levels, nlev, levels[0].n);
  if (rc) return rc;
  SMX_REQUIRE(r && u && work, SMX_ERR_VALUE, "null pointer argument");
  hipStream_t st = smx::as_stream(stream);
  return val_dtype == SMX_F64 ? apply_once<double>(levels, nlev, r, u, work, st)
                              : apply_once<float>(levels, nlev, r, u, work, st);
}

extern "C" int64_t smx_pcg_amg_ws_size(int64_t n) {
  const Layout L = layout(n, n);
  return HDR + 7 * L.Gf + 9 * L.nvec + 64;
}

namespace {

int amg_entry(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols, const double* vals,
              int32_t mode, const double* b, const double* b_part, int64_t nb_part,
              int32_t project, int32_t raw_width, int64_t raw_ld, const int32_t* raw_cols,
              const double* raw_vals, const int32_t* bpos, const double* gb, double* x,
              int32_t warm, const int32_t* perm, const smx_amg_level* levels, int32_t nlev,
              int32_t val_dtype, double rtol, double atol, int32_t max_it, int32_t check_every,
              int32_t expected_its, double* work, int32_t* its_out, double* rnorm_out,
              double* sums_out, const smx_cg_mixed* mx, void* stream) {
  SMX_REQUIRE(n >= 0 && nd >= 0, SMX_ERR_SHAPE, "negative size");
  SMX_REQUIRE(mode == 0 || mode == 1, SMX_ERR_VALUE, "mode must be 0 or 1");
  SMX_REQUIRE(nd == 0 || mode == 1 || !project, SMX_ERR_VALUE,
              "trailing identity rows (nd > 0) in mode 0 need project = 0");
  SMX_REQUIRE(max_it >= 0, SMX_ERR_VALUE, "max_it must be >= 0");
  SMX_REQUIRE(check_every >= 0 || (expected_its > 0 && check_every >= -64), SMX_ERR_VALUE,
              "a speculative launch (check_every in [-64, -1]) needs expected_its > 0");
  SMX_REQUIRE(width >= 1 && width <= 32 && ld >= n, SMX_ERR_SHAPE,
              "ELL width must be in [1, 32] and ld >= n (width %d, ld %lld)", width, (long long)ld);
  SMX_REQUIRE(mode == 0 || (raw_width >= 1 && raw_width <= 32 && raw_ld >= n), SMX_ERR_SHAPE,
              "raw operator width must be in [1, 32]");
  SMX_REQUIRE(val_dtype == SMX_F64 || val_dtype == SMX_F32, SMX_ERR_TYPE,
              "AMG operator values must be FP64 or FP32");
  if (mx) {
    SMX_REQUIRE(mx->vals32, SMX_ERR_VALUE, "mixed precision needs the FP32 operator values");
    SMX_REQUIRE(mx->refinements >= 0 && mx->refinements <= 8, SMX_ERR_VALUE,
                "refinements must be in [0, 8]");
    SMX_REQUIRE(mx->rtol_inner > 0.0 && mx->rtol_inner < 1.0, SMX_ERR_VALUE,
                "rtol_inner must be in (0, 1)");
  }
  if (its_out) *its_out = 0;
  if (rnorm_out) *rnorm_out = 0.0;
  SMX_REQUIRE(n > 0 || nd == 0, SMX_ERR_SHAPE, "trailing identity rows need n >= 1");
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(cols && vals && x && work, SMX_ERR_VALUE, "null pointer argument");
  SMX_REQUIRE(mode == 1 || b, SMX_ERR_VALUE, "mode 0 needs b");
  SMX_REQUIRE(mode == 0 || (raw_cols && raw_vals && bpos && gb), SMX_ERR_VALUE,
              "mode 1 needs the raw operator, bpos and gb");
  {
    const int rc = check_levels(levels, nlev, n);
    if (rc) return rc;
  }
  SMX_REQUIRE(n + nd < (int64_t)1 << 30, SMX_ERR_SHAPE, "operator too large (n %lld)",
              (long long)(n + nd));
  hipStream_t st = smx::as_stream(stream);
  const Layout Lo = layout(n, n + nd);
  AmgPtrs P;
  P.Gf = (int)Lo.Gf;
  P.ss = work;
  P.st[0] = work + ST_OFF;
  P.st[1] = work + ST_OFF + ST_SIZE;
  double* q = work + HDR;
  P.pbb = q;
  P.prr[0] = q + Lo.Gf;
  P.prr[1] = q + 3 * Lo.Gf;
  P.pd = q + 5 * Lo.Gf;
  P.pg = q + 6 * Lo.Gf;
  double* v0 = q + 7 * Lo.Gf;
  v0 += (64 - (((uintptr_t)v0 / sizeof(double)) & 63)) & 63;  // 512-byte aligned vectors
  P.r[0] = v0;
  P.r[1] = v0 + Lo.nvec;
  P.s[0] = v0 + 2 * Lo.nvec;
  P.s[1] = v0 + 3 * Lo.nvec;
  P.w = v0 + 4 * Lo.nvec;
  P.u = v0 + 5 * Lo.nvec;
  P.p = v0 + 6 * Lo.nvec;
  double* x_perm = v0 + 7 * Lo.nvec;
  P.z1 = v0 + 8 * Lo.nvec;  // n_1 < n
  double hst[SS_SIZE] = {0};
  const int bs = fine_bs(n);
  const int32_t wmax = (mode == 1 && raw_width > width) ? raw_width : width;
  const bool w7 = width == 7 && wmax == 7;
  const int32_t nref = mx ? mx->refinements : 0;
  const double rin = mx ? mx->rtol_inner : rtol;
  int rc;
#define SMX_AMG_CALL(BS, WM, EX, VT, V, AV, CP, IV)                                              \
  run_amg<BS, WM, EX, VT, V, AV, CP>(n, nd, width, ld, cols, vals, IV, mode, b, b_part, nb_part, \
                                     project, raw_width, raw_ld, raw_cols, raw_vals, bpos, gb, x, \
                                     warm, perm, x_perm, levels, nlev, rtol, atol, max_it,       \
                                     check_every, expected_its, nref, rin, P, st, hst)
#define SMX_AMG_W(BS, VT, V, AV, CP, IV)                                                          \
  (w7 ? SMX_AMG_CALL(BS, 7, true, VT, V, AV, CP, IV)                                              \
      : wmax <= 8 ? SMX_AMG_CALL(BS, 8, false, VT, V, AV, CP, IV)                                 \
                  : wmax <= 16 ? SMX_AMG_CALL(BS, 16, false, VT, V, AV, CP, IV)                   \
                               : SMX_AMG_CALL(BS, 32, false, VT, V, AV, CP, IV))
  if (mx) {
    const float* v32 = mx->vals32;
    if (val_dtype == SMX_F64) {
      rc = bs == 256 ? SMX_AMG_W(256, double, float, float, true, v32)
                     : SMX_AMG_W(512, double, float, float, true, v32);
    } else {
      rc = bs == 256 ? SMX_AMG_W(256, float, float, float, true, v32)
                     : SMX_AMG_W(512, float, float, float, true, v32);
    }
  } else if (val_dtype == SMX_F64) {
    rc = bs == 256 ? SMX_AMG_W(256, double, double, double, false, vals)
                   : SMX_AMG_W(512, double, double, double, false, vals);
  } else {
    rc = bs == 256 ? SMX_AMG_W(256, float, double, double, false, vals)
                   : SMX_AMG_W(512, float, double, double, false, vals);
  }
#undef SMX_AMG_W
#undef SMX_AMG_CALL
  if (rc != SMX_OK) return rc;
  if (check_every < 0) {  // speculative: nothing waited for, the status stays on the device
    if (its_out) *its_out = -1;
    return SMX_OK;
  }
  if (its_out) *its_out = (int32_t)(hst[SS_ITERS] + hst[SS_ITERS0]);
  if (rnorm_out) *rnorm_out = sqrt(hst[SS_RR]);
  if (sums_out) sums_out[0] = hst[SS_SUMX];
  if (hst[SS_FAIL] == 1.0) {
    smx::set_error("AMG-CG breakdown after %d iterations (rnorm %g)", (int)hst[SS_ITERS],
                   sqrt(hst[SS_RR]));
    return SMX_ERR_NOT_CONVERGED;
  }
  if (hst[SS_FAIL] == 2.0 || hst[SS_DONE] == 0.0) {
    smx::set_error("AMG-CG did not converge in %d iterations (rnorm %g, tol %g)", max_it,
                   sqrt(hst[SS_RR]), sqrt(hst[SS_TOL2]));
    return SMX_ERR_NOT_CONVERGED;
  }
  return SMX_OK;
}

}  // namespace

extern "C" int smx_pcg_amg(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols,
                           const double* vals, int32_t mode, const double* b,
                           const double* b_part, int64_t nb_part, int32_t project,
                           int32_t raw_width, int64_t raw_ld, const int32_t* raw_cols,
                           const double* raw_vals, const int32_t* bpos, const double* gb,
                           double* x, int32_t warm, const int32_t* perm,
                           const smx_amg_level* levels, int32_t nlev, int32_t val_dtype,
                           double rtol, double atol, int32_t max_it, int32_t check_every,
                           int32_t expected_its, double* work, int32_t* its_out,
                           double* rnorm_out, double* sums_out, void* stream) {
  return amg_entry(n, nd, width, ld, cols, vals, mode, b, b_part, nb_part, project, raw_width, raw_ld,
                   raw_cols, raw_vals, bpos, gb, x, warm, perm, levels, nlev, val_dtype, rtol, atol,
                   max_it, check_every, expected_its, work, its_out, rnorm_out, sums_out, nullptr,
                   stream);
}

extern "C" int smx_pcg_amg_mixed(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols,
                                 const double* vals, int32_t mode, const double* b,
                                 const double* b_part, int64_t nb_part, int32_t project,
                                 int32_t raw_width, int64_t raw_ld, const int32_t* raw_cols,
                                 const double* raw_vals, const int32_t* bpos, const double* gb,
                                 double* x, int32_t warm, const int32_t* perm,
                                 const smx_amg_level* levels, int32_t nlev, int32_t val_dtype,
                                 double rtol, double atol, int32_t max_it, int32_t check_every,
                                 int32_t expected_its, double* work, int32_t* its_out,
                                 double* rnorm_out, double* sums_out, const smx_cg_mixed* mixed,
                                 void* stream) {
  SMX_REQUIRE(mixed, SMX_ERR_VALUE, "null mixed-precision options");
  return amg_entry(n, nd, width, ld, cols, vals, mode, b, b_part, nb_part, project, raw_width, raw_ld,
                   raw_cols, raw_vals, bpos, gb, x, warm, perm, levels, nlev, val_dtype, rtol, atol,
                   max_it, check_every, expected_its, work, its_out, rnorm_out, sums_out, mixed,
                   stream);
}
