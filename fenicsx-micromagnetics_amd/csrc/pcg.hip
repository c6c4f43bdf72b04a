// Jacobi-preconditioned conjugate gradients for the two FEM Poisson solves.
//
// Chronopoulos-Gear formulation: both inner products of an iteration come
// from one reduction, so an iteration is ONE kernel launch:
//   s_i = w_i + b_i s_{i-1};  p_i = u_i + b_i p_{i-1};  x += a_i p_i
//   r_{i+1} = r_i - a_i s_i;  u = D^-1 r;  w_{i+1} = A u_{i+1}
//   g = (r,u), d = (w,u), rr = (r,r);  b = g/g_old;  a = g/(d - b g/a_old)
// The SpMV needs u_{i+1} at neighbour rows, which each thread rebuilds from
// the previous iterate (r, w, s are double-buffered), so no grid-wide barrier
// is needed inside an iteration.
//
// Reductions without atomics or device-scope fences: every workgroup writes
// its four partial sums; the NEXT launch's workgroups each re-reduce all of
// them in the same fixed order (the kernel boundary is the only
// synchronization), so every workgroup derives identical scalars.  A
// device-scope release per workgroup (an L2 write-back on a multi-XCD part)
// costs more than re-reading a few KB of partials from the Infinity Cache.
// Scalars carried between launches are double-buffered by iteration parity;
// workgroup 0 also publishes a status record that the host polls every
// `check_every` launches.  Once converged every later launch returns early.
//
// Matrices are ELL: column j of row i at cols[j*ld + i] (width W, padding
// entries have value 0 and a valid column), so a wavefront's index and value
// loads are contiguous and the row start needs no extra dependent load.
// The CG operator is given in SOLVER FORM (smx_pcg_prepare builds it): the
// diagonal entry of every row sits in slot 0 and every value is a_ij * dinv_j,
// i.e. the matrix is A D^-1.  Then
//   w_{i+1} = A D^-1 r_{i+1} = sum_j (a_ij dinv_j) (r_j - a s_j)
// needs three neighbour values per entry instead of four, and the row's own
// r, w, s come from the slot-0 gather instead of separate loads.
//
// Cost model (film operator, N = 202005, 7 entries per row): a dependent
// empty launch costs about 2.8 us and the 27 MB an iteration streams add about
// 6 us, so an iteration is bound by the kernel boundary and the memory system
// together, not by instruction issue.  What sits on the dependent chain inside
// a launch is kept short:
//  - the partial sums are the FIRST loads issued, and every load before the
//    reduction is unconditional (rows past n read row n-1), so the reduction
//    waits on the partials only while the streaming loads are in flight (the
//    hardware retires loads in issue order, and a branch around some loads
//    makes the compiler wait as if they had not been issued);
//  - each reduction is one DPP wave total plus ONE LDS-only barrier;
//  - loads use a uniform base plus a 32-bit offset, one offset register per
//    neighbour shared by r, w and s;
//  - a stencil of width 7 is compiled with its exact width.
// Gathering the neighbour values before the first reduction needs more than
// 128 VGPRs, which halves the resident workgroups per CU and loses.
//
// Mixed precision (smx_pcg_jacobi_mixed): the iteration vectors r, w, s, p
// and the iteration operator are FP32, halving the bytes an iteration
// streams; x, the scalars and every inner product stay FP64.  A product of
// two FP32 numbers is exact in FP64, and the sums are compensated (TwoSum in
// the wave butterfly and across workgroups), so the inner products carry no
// more than a few FP64 roundings.  Accuracy is recovered by iterative
// refinement: after the FP32 solve, the residual b - A x is recomputed in FP64
// from the FP64 operator and checked against the tolerance (two launches);
// only when it fails does a correction solve run, from the current x.
#include "smx_common.h"

namespace {

enum {
  // status record (written by workgroup 0, read by the host)
  SS_DONE = 0,
  SS_ITERS = 1,
  SS_RR = 2,
  SS_TOL2 = 3,
  SS_BB = 4,
  SS_FAIL = 5,  // 1 breakdown, 2 max_it reached, 3 operator not in solver form
  SS_SUMX = 6,
  SS_ITERS0 = 7,  // iterations of the inner solves before the last (mixed precision)
  SS_SIZE = 8,
  // scalars carried to the next launch: [gamma, alpha, 1/gamma, 1/alpha]
  ST_OFF = 8,
  ST_SIZE = 8,
  HDR = 32
};
static_assert(SS_DONE == SMX_PCG_DONE && SS_ITERS == SMX_PCG_ITERS && SS_FAIL == SMX_PCG_FAIL &&
                  SS_SUMX == SMX_PCG_SUMX,
              "status block layout of the C ABI");

// Iteration vectors are FP64 or FP32 (the element type is a kernel template
// parameter); x is always FP64.
struct CGPtrs {
  double* ss;        // status record
  double* st[2];     // parity-buffered scalars
  double* part[2];   // parity-buffered partials, 4 x G
  double* pbb;       // 3 x G partials: ||b||^2, rows whose slot 0 is not the diagonal, ||b - A x0||^2
  void* r[2];
  void* w[2];
  void* s[2];
  void* p;
  int G;
};

template <typename V>
__host__ __device__ __forceinline__ V* vec(void* p) {
  return static_cast<V*>(p);
}


// Hardware dispatch places workgroup b on XCD b % 8.  Renumbering so that each
// XCD processes one contiguous range of rows keeps most neighbour rows of the
// SpMV inside one L2 (bijective for every G).
__device__ __forceinline__ int xcd_block(int G) {
  const int b = blockIdx.x, x = b & 7, k = b >> 3;
  const int q = G >> 3, r = G & 7;
  return x * q + (x < r ? x : r) + k;
}

// Element at a byte offset from a uniform base: the SGPR-base + VGPR-offset
// load form, without a 64-bit address register per load.
template <typename T>
__device__ __forceinline__ T ldo(const T* __restrict__ base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Workgroup sum of NQ values per thread.  Wave totals go through DPP lane
// moves; the waves meet at ONE LDS-only barrier, so each call site needs its
// own `scr` (NQ * BS/64 doubles): a buffer reused by a later reduction in the
// same launch would need a second barrier.  ALL: every thread gets the totals
// (else only thread 0).  The summation order is fixed, so every workgroup that
// reduces the same values gets bit-identical totals.
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

// Sum of NQ x G partials (quantity q of block b at part[q*G + b]), identical
// in every workgroup and every thread.
template <int BS, int NQ, bool COMP = false>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int G,
                                                double (&t)[NQ], double* scr) {
  double c[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) t[q] = c[q] = 0.0;
  for (int b = threadIdx.x; b < G; b += BS) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (COMP) {
        smx::comp_add(t[q], c[q], part[q * G + b]);
      } else {
        t[q] += part[q * G + b];
      }
    }
  }
  block_total<BS, NQ, true, COMP>(t, scr, c);
}

template <int BS, int NQ, bool COMP = false>
__device__ __forceinline__ void publish_partials(double (&v)[NQ], double* __restrict__ part, int G,
                                                 double* scr, int blk = -1) {
  if (blk < 0) blk = blockIdx.x;
  block_total<BS, NQ, false, COMP>(v, scr);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) part[q * G + blk] = v[q];
  }
}

// Solver form, one thread per row: the first slot whose column is the row
// moves to slot 0 (slot 0's entry takes its place) and every value is scaled
// by dinv of its column.  A row without a diagonal entry is copied unchanged
// (the solver's layout check then rejects the operator).  out may alias in:
// every slot is read before it is written, slots 0 and d from saved copies.
__global__ __launch_bounds__(256) void cg_prepare_kernel(int64_t n, int32_t W, int64_t ld,
                                                         const int32_t* cols, const double* vals,
                                                         const double* __restrict__ dinv,
                                                         int32_t* cols_out, double* vals_out) {
  const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  int d = -1;
  for (int j = 0; j < W; ++j) {
    if (cols[j * ld + row] == (int32_t)row) {
      d = j;
      break;
    }
  }
  if (d < 0) {
    for (int j = 0; j < W; ++j) {
      cols_out[j * ld + row] = cols[j * ld + row];
      vals_out[j * ld + row] = vals[j * ld + row];
    }
    return;
  }
  const int32_t c0 = cols[row], cd = cols[d * ld + row];
  const double v0 = vals[row], vd = vals[d * ld + row];
  for (int j = 0; j < W; ++j) {
    const int32_t c = j == 0 ? cd : (j == d ? c0 : cols[j * ld + row]);
    const double v = j == 0 ? vd : (j == d ? v0 : vals[j * ld + row]);
    cols_out[j * ld + row] = c;
    vals_out[j * ld + row] = v * dinv[c];
  }
}

// r0 = b - A x0 (mode 0) or the lifted residual (mode 1), in FP64 from the
// FP64 operator; partials of ||b||^2, of the rows whose slot 0 is not their
// diagonal (acols: the solver-form operator, checked in both modes) and of
// ||r0||^2 (read by the refinement check).
template <int BS, int WMAX, typename V>
__global__ __launch_bounds__(BS) void cg_init_a_kernel(
    int64_t n, int64_t nd, int32_t mode, int32_t W, int64_t ld, const int32_t* __restrict__ cols,
    const double* __restrict__ vals, const int32_t* __restrict__ acols,
    const double* __restrict__ dinv, const double* __restrict__ b,
    const double* __restrict__ b_part, int64_t nb_part, int32_t project,
    const int32_t* __restrict__ bpos, const double* __restrict__ gb, double* __restrict__ x,
    int32_t warm, const int32_t* __restrict__ perm, CGPtrs P) {
  __shared__ double scr[BS / 64 * 1];
  __shared__ double scr_pub[BS / 64 * 3];
  const int64_t row = (int64_t)blockIdx.x * BS + threadIdx.x;
  double shift = 0.0;
  if (mode == 0 && project && b_part) {
    // sum of the right-hand side partials (the divergence kernel's blocks)
    double t[1] = {0.0};
    for (int64_t k = threadIdx.x; k < nb_part; k += BS) t[0] += b_part[k];
    block_total<BS, 1, true>(t, scr);
    shift = t[0] / (double)n;
  }
  double v[3] = {0.0, 0.0, 0.0};
  if (row < n) {
    double r0, bk;
    if (mode == 0) {
      bk = b[perm ? perm[row] : row] - shift;
      double ax = 0.0;
      if (warm) {
        // solver-form values are a_ij dinv_j: A x = sum_j (a_ij dinv_j) (x_j / dinv_j)
#pragma unroll
        for (int j = 0; j < WMAX; ++j) {
          if (j < W) {
            const int32_t c = cols[j * ld + row];
            ax = fma(vals[j * ld + row], x[c] / dinv[c], ax);
          }
        }
      } else {
        x[row] = 0.0;
      }
      r0 = bk - ax;
    } else {
      const int32_t bk_pos = bpos[row];
      if (bk_pos >= 0) {
        bk = gb[bk_pos];
        x[row] = bk;
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
        if (!warm) x[row] = 0.0;
        r0 = -sx;
        bk = -sg;
      }
    }
    vec<V>(P.r[0])[row] = (V)r0;
    vec<V>(P.s[0])[row] = (V)0.0;
    vec<V>(P.p)[row] = (V)0.0;
    v[0] = bk * bk;
    v[1] = acols[row] != (int32_t)row ? 1.0 : 0.0;
    v[2] = r0 * r0;
  }
  // trailing identity rows n .. n+nd-1: x is their data (b, or gb at bpos),
  // their residual is 0 and they count in ||b|| as rows of the full system
  for (int64_t t = n + row; t < n + nd; t += (int64_t)gridDim.x * BS) {
    const double bt = mode == 0 ? b[perm ? perm[t] : t] : gb[bpos[t]];
    x[t] = bt;
    v[0] = fma(bt, bt, v[0]);
  }
  publish_partials<BS, 3>(v, P.pbb, P.G, scr_pub);
}

// Refinement check, one workgroup: the FP64 residual ||b - A x|| left by
// cg_init_a against the tolerance.  Met: the solve is done (x stands).  Not
// met: the iterations so far move to SS_ITERS0 and the status is reset for a
// correction solve from the current x.
template <int BS>
__global__ __launch_bounds__(BS) void cg_check_kernel(CGPtrs P, double rtol, double atol) {
  __shared__ double scr[BS / 64 * 3];
  double t[3];
  reduce_partials<BS, 3, true>(P.pbb, P.G, t, scr);
  if (threadIdx.x == 0) {
    const double tol2 = fmax(rtol * rtol * t[0], atol * atol);
    const bool ok = t[2] <= tol2 && t[1] == 0.0;
    P.ss[SS_BB] = t[0];
    P.ss[SS_TOL2] = tol2;
    P.ss[SS_RR] = t[2];
    P.ss[SS_FAIL] = t[1] > 0.0 ? 3.0 : 0.0;
    if (ok || t[1] > 0.0) {
      P.ss[SS_DONE] = 1.0;
    } else {
      P.ss[SS_ITERS0] = P.ss[SS_ITERS0] + P.ss[SS_ITERS];
      P.ss[SS_ITERS] = 0.0;
      P.ss[SS_DONE] = 0.0;
    }
  }
}

// x_p[i] = x[perm[i]] (warm start in solver order)
__global__ void cg_gather_kernel(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ x,
                                 double* __restrict__ xp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xp[i] = x[perm[i]];
}

// x[perm[i]] = x_p[i] (solution back in the caller's order)
__global__ void cg_scatter_kernel(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ xp,
                                  double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[perm[i]] = xp[i];
}

// w0 = A D^-1 r0 with the iteration operator (values AV); partials of
// (g0, d0, rr0, sum x0).  rtol: the relative tolerance of this solve; phase 0
// clears SS_ITERS0 (a correction solve keeps what the check moved there).
template <int BS, int WMAX, typename V, typename AV, bool COMP>
__global__ __launch_bounds__(BS) void cg_init_b_kernel(int64_t n, int32_t W, int64_t ld,
                                                       const int32_t* __restrict__ cols,
                                                       const AV* __restrict__ vals,
                                                       const double* __restrict__ dinv,
                                                       const double* __restrict__ x, CGPtrs P,
                                                       double rtol, double atol, int32_t phase) {
  __shared__ double scr[BS / 64 * 2];
  __shared__ double scr_pub[BS / 64 * 4];
  double t[2];
  reduce_partials<BS, 2>(P.pbb, P.G, t, scr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double t1 = rtol * rtol * t[0], t2 = atol * atol;
    P.ss[SS_BB] = t[0];
    P.ss[SS_TOL2] = t1 > t2 ? t1 : t2;
    if (phase == 0) P.ss[SS_ITERS0] = 0.0;
    P.ss[SS_ITERS] = 0.0;
    P.ss[SS_DONE] = t[1] > 0.0 ? 1.0 : 0.0;  // a layout error stops before iteration 0
    P.ss[SS_FAIL] = t[1] > 0.0 ? 3.0 : 0.0;
  }
  const int64_t row = (int64_t)blockIdx.x * BS + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (row < n) {
    const V* r = vec<V>(P.r[0]);
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < WMAX; ++j) {
      if (j < W) acc = fma((double)vals[j * ld + row], (double)r[cols[j * ld + row]], acc);
    }
    const V w0 = (V)acc;
    vec<V>(P.w[0])[row] = w0;
    const double rk = (double)r[row];
    const double uk = rk * dinv[row];
    v[0] = rk * uk;
    v[1] = (double)w0 * uk;
    v[2] = rk * rk;
    v[3] = x[row];
  }
  publish_partials<BS, 4, COMP>(v, P.part[0], P.G, scr_pub);
}

// One CG iteration.  EXACT: the ELL width is WMAX (no per-slot predicate).
// V: iteration vectors, AV: iteration operator values (FP64 or FP32); the
// arithmetic is FP64 throughout.
template <int BS, int WMAX, bool EXACT, typename V, typename AV, bool COMP>
__global__ __launch_bounds__(BS) void cg_iter_kernel(int64_t n, int32_t W, int64_t ld,
                                                     const int32_t* __restrict__ cols,
                                                     const AV* __restrict__ vals,
                                                     const double* __restrict__ dinv,
                                                     double* __restrict__ x, CGPtrs P, int32_t it,
                                                     int32_t max_it) {
  __shared__ double scr[BS / 64 * 4];
  __shared__ double scr_pub[BS / 64 * 4];
  // Written by an earlier launch; tested once the loads below are in flight.
  const bool was_done = P.ss[SS_DONE] != 0.0;
  const int q = it & 1;
  const V* __restrict__ rq = vec<V>(P.r[q]);
  const V* __restrict__ wq = vec<V>(P.w[q]);
  const V* __restrict__ sq = vec<V>(P.s[q]);
  const int G = P.G;
  const double* __restrict__ part = P.part[q];
  const int b0 = (int)threadIdx.x < G ? (int)threadIdx.x : 0;
  double t[4], tc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; ++k) t[k] = part[k * G + b0];
  const int blk = xcd_block(G);
  const int64_t row = (int64_t)blk * BS + threadIdx.x;
  const int64_t rowc = row < n ? row : n - 1;
  const int Wk = EXACT ? WMAX : W;
  constexpr uint32_t SV = sizeof(V), SA = sizeof(AV);
  const uint32_t o4 = (uint32_t)rowc * 4u, o8 = (uint32_t)rowc * 8u;
  const uint32_t oa = (uint32_t)rowc * SA, ov = (uint32_t)rowc * SV;
  int32_t c[WMAX];  // byte offsets are formed after the reduction, where they are used
  double av[WMAX];
#pragma unroll
  for (int j = 0; j < WMAX; ++j) {
    c[j] = j < Wk ? ldo(cols + j * ld, o4) : (int32_t)rowc;
    av[j] = j < Wk ? (double)ldo(vals + j * ld, oa) : 0.0;
  }
  const double dk = ldo(dinv, o8);
  const double p0 = (double)ldo((const V*)P.p, ov);
  const double x0 = ldo((const double*)x, o8);
  if ((int)threadIdx.x >= G) {
#pragma unroll
    for (int k = 0; k < 4; ++k) t[k] = 0.0;
  }
  for (int b = threadIdx.x + BS; b < G; b += BS) {  // only when G > BS
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (COMP) {
        smx::comp_add(t[k], tc[k], part[k * G + b]);
      } else {
        t[k] += part[k * G + b];
      }
    }
  }
  block_total<BS, 4, true, COMP>(t, scr, tc);
  const double gam = t[0], del = t[1], rr = t[2];
  const double tol2 = P.ss[SS_TOL2];
  double alpha, beta;
  if (it == 0) {
    beta = 0.0;
    alpha = gam / del;
  } else {
    // reciprocals stored by the previous launch: one division on the chain
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
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (row < n) {
    double rj[WMAX], wj[WMAX], sj[WMAX];
#pragma unroll
    for (int j = 0; j < WMAX; ++j) {
      const uint32_t oc = (uint32_t)c[j] * SV;
      rj[j] = (double)ldo(rq, oc);
      wj[j] = (double)ldo(wq, oc);
      sj[j] = (double)ldo(sq, oc);
    }
    // neighbour values of r_{i+1}, by the same explicit fmas and roundings as
    // the row's own r_n below, so w_{i+1} is A D^-1 applied to the stored r
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < WMAX; ++j) {
      const double sn_j = (double)(V)fma(beta, sj[j], wj[j]);
      acc = fma(av[j], (double)(V)fma(-alpha, sn_j, rj[j]), acc);
    }
    const double r0 = rj[0], w0 = wj[0], s0 = sj[0];  // slot 0 is the row itself
    const V s_n = (V)fma(beta, s0, w0);
    const V p_n = (V)fma(beta, p0, r0 * dk);
    const double x_n = fma(alpha, (double)p_n, x0);
    const V r_n = (V)fma(-alpha, (double)s_n, r0);
    const V w_n = (V)acc;
    const double u_n = (double)r_n * dk;
    // streaming stores: the next launch reads these from another XCD anyway,
    // and lines that bypass L2 need no write-back at the kernel boundary
    __builtin_nontemporal_store(s_n, &vec<V>(P.s[1 - q])[row]);
    __builtin_nontemporal_store(r_n, &vec<V>(P.r[1 - q])[row]);
    __builtin_nontemporal_store(w_n, &vec<V>(P.w[1 - q])[row]);
    __builtin_nontemporal_store(p_n, &vec<V>(P.p)[row]);
    __builtin_nontemporal_store(x_n, &x[row]);
    v[0] = (double)r_n * u_n;
    v[1] = (double)w_n * u_n;
    v[2] = (double)r_n * (double)r_n;
    v[3] = x_n;
  }
  publish_partials<BS, 4, COMP>(v, P.part[1 - q], G, scr_pub, blk);
}

int block_size_for(int64_t n) { return n <= 65536 ? 256 : 512; }

int64_t grid_for_n(int64_t n) {
  const int bs = block_size_for(n);
  return (n + bs - 1) / bs;
}

struct Layout {
  int64_t G, nvec;
};

// G from the n iterated rows; vectors (x in solver order among them) hold
// ntot = n + nd rows, so smx_pcg_ws_size(ntot) covers the layout
Layout layout(int64_t n, int64_t ntot) {
  Layout L;
  L.G = grid_for_n(n);
  L.nvec = (ntot + 63) / 64 * 64;
  return L;
}

// The solve: CG from x0 (warm) or 0 to rtol_inner (rtol without
// refinement; rtol_inner may be below rtol), then up to `refinements` times: the FP64 residual b - A x and
// its check (cg_init_a + cg_check_kernel); if it fails, a correction solve
// from the current x to the next tolerance.  A speculative launch enqueues
// the first solve and one check and returns: a failed check leaves DONE = 0
// in the status block, and the caller solves again waiting.
template <int BS, int WMAX, bool EXACT, typename V, typename AV, bool COMP>
int run_pcg(int64_t n, int64_t nd, int32_t W, int64_t ld, const int32_t* cols, const double* vals,
            const AV* ivals, const double* dinv, int32_t mode, const double* b,
            const double* b_part, int64_t nb_part, int32_t project, int32_t raw_W, int64_t raw_ld,
            const int32_t* raw_cols, const double* raw_vals, const int32_t* bpos, const double* gb,
            double* x_user, int32_t warm, const int32_t* perm, double* x_perm, double rtol,
            double atol, int32_t max_it, int32_t check_every, int32_t expected_its,
            int32_t refinements, double rtol_inner, CGPtrs P, hipStream_t st, double* hst) {
  const int G = P.G;
  const unsigned gs = (unsigned)((n + nd + 255) / 256);
  double* x = perm ? x_perm : x_user;  // the kernels work in solver (permuted) row order
  if (perm && warm) {
    hipLaunchKernelGGL(cg_gather_kernel, dim3(gs), dim3(256), 0, st, n + nd, perm, x_user, x);
    SMX_LAUNCH_CHECK();
  }
  // check_every < 0: speculative launch of expected + 1 - check_every
  // iterations with no wait (iterations past convergence return at once on
  // the done flag); the caller reads the status block later
  const bool spec = check_every < 0 && expected_its > 0;
  const int32_t every = check_every < 1 ? 16 : check_every;
  auto init_a = [&](int32_t w) -> int {
    if (mode == 0) {
      hipLaunchKernelGGL((cg_init_a_kernel<BS, WMAX, V>), dim3(G), dim3(BS), 0, st, n, nd, mode, W, ld,
                         cols, vals, cols, dinv, b, b_part, nb_part, project, bpos, gb, x, w, perm, P);
    } else {
      hipLaunchKernelGGL((cg_init_a_kernel<BS, WMAX, V>), dim3(G), dim3(BS), 0, st, n, nd, mode, raw_W,
                         raw_ld, raw_cols, raw_vals, cols, dinv, b, b_part, nb_part, project, bpos, gb,
                         x, w, perm, P);
    }
    SMX_LAUNCH_CHECK();
    return SMX_OK;
  };
  // one CG solve on the residual in P.r[0] (left by cg_init_a)
  auto iterate = [&](int32_t phase, double tol, int32_t expect) -> int {
    hipLaunchKernelGGL((cg_init_b_kernel<BS, WMAX, V, AV, COMP>), dim3(G), dim3(BS), 0, st, n, W,
                       ld, cols, ivals, dinv, x, P, tol, atol, phase);
    SMX_LAUNCH_CHECK();
    int32_t batch = spec ? expect + 1 - check_every : (expect > 0 ? expect + 1 : every);
    int32_t launched = 0;  // iterations launched: it = 0 .. launched-1
    bool done = false;
    while (!done && launched <= max_it) {
      const int32_t nb = batch < (max_it + 1 - launched) ? batch : (max_it + 1 - launched);
      for (int32_t k = 0; k < nb; ++k) {
        hipLaunchKernelGGL((cg_iter_kernel<BS, WMAX, EXACT, V, AV, COMP>), dim3(G), dim3(BS), 0,
                           st, n, W, ld, cols, ivals, dinv, x, P, launched + k, max_it);
      }
      SMX_LAUNCH_CHECK();
      launched += nb;
      if (spec) break;
      SMX_HIP_CHECK(hipMemcpyAsync(hst, P.ss, sizeof(double) * SS_SIZE, hipMemcpyDeviceToHost, st));
      SMX_HIP_CHECK(hipStreamSynchronize(st));
      done = hst[SS_DONE] != 0.0;
      batch = every;
    }
    return SMX_OK;
  };
  int rc = init_a(warm);
  if (rc) return rc;
  rc = iterate(0, refinements > 0 ? rtol_inner : rtol, expected_its);
  if (rc) return rc;
  for (int32_t k = 1; k <= refinements; ++k) {
    rc = init_a(1);
    if (rc) return rc;
    hipLaunchKernelGGL((cg_check_kernel<1024>), dim3(1), dim3(1024), 0, st, P, rtol, atol);
    SMX_LAUNCH_CHECK();
    if (spec) break;
    SMX_HIP_CHECK(hipMemcpyAsync(hst, P.ss, sizeof(double) * SS_SIZE, hipMemcpyDeviceToHost, st));
    SMX_HIP_CHECK(hipStreamSynchronize(st));
    if (hst[SS_DONE] != 0.0) break;
    // correction k stops at rtol_inner^(k+1) of ||b||, the last at rtol
    const double tol = k == refinements ? rtol : fmax(rtol, pow(rtol_inner, (double)(k + 1)));
    rc = iterate(k, tol, 0);
    if (rc) return rc;
  }
  if (perm) {
    hipLaunchKernelGGL(cg_scatter_kernel, dim3(gs), dim3(256), 0, st, n + nd, perm, x, x_user);
    SMX_LAUNCH_CHECK();
  }
  if (!spec) {
    SMX_HIP_CHECK(hipMemcpyAsync(hst, P.ss, sizeof(double) * SS_SIZE, hipMemcpyDeviceToHost, st));
    SMX_HIP_CHECK(hipStreamSynchronize(st));
  }
  return SMX_OK;
}

}  // namespace

extern "C" int smx_pcg_prepare(int64_t n, int32_t width, int64_t ld, const int32_t* cols,
                               const double* vals, const double* dinv, int32_t* cols_out,
                               double* vals_out, void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative size");
  SMX_REQUIRE(width >= 1 && width <= 32 && ld >= n, SMX_ERR_SHAPE,
              "ELL width must be in [1, 32] and ld >= n (width %d, ld %lld)", width, (long long)ld);
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(cols && vals && dinv && cols_out && vals_out, SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(cg_prepare_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     smx::as_stream(stream), n, width, ld, cols, vals, dinv, cols_out, vals_out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int64_t smx_pcg_ws_size(int64_t n) {
  const Layout L = layout(n, n);
  return HDR + 11 * L.G + 8 * L.nvec + 64;
}

namespace {

int pcg_entry(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols, const double* vals,
              const double* dinv, int32_t mode, const double* b, const double* b_part,
              int64_t nb_part, int32_t project, int32_t raw_width, int64_t raw_ld,
              const int32_t* raw_cols, const double* raw_vals, const int32_t* bpos,
              const double* gb, double* x, int32_t warm, const int32_t* perm, double rtol,
              double atol, int32_t max_it, int32_t check_every, int32_t expected_its,
              double* work, int32_t* its_out, double* rnorm_out, double* sums_out,
              const smx_cg_mixed* mx, void* stream) {
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
  SMX_REQUIRE(n + nd < (int64_t)1 << 28, SMX_ERR_SHAPE,
              "CG operator too large for 32-bit byte offsets (n %lld)", (long long)(n + nd));
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
  SMX_REQUIRE(cols && vals && dinv && x && work, SMX_ERR_VALUE, "null pointer argument");
  SMX_REQUIRE(mode == 1 || b, SMX_ERR_VALUE, "mode 0 needs b");
  SMX_REQUIRE(mode == 0 || (raw_cols && raw_vals && bpos && gb), SMX_ERR_VALUE,
              "mode 1 needs the raw operator, bpos and gb");
  hipStream_t st = smx::as_stream(stream);
  const Layout Lo = layout(n, n + nd);
  CGPtrs P;
  P.G = (int)Lo.G;
  P.ss = work;
  P.st[0] = work + ST_OFF;
  P.st[1] = work + ST_OFF + ST_SIZE;
  P.part[0] = work + HDR;
  P.part[1] = work + HDR + 4 * Lo.G;
  P.pbb = work + HDR + 8 * Lo.G;
  double* v0 = work + HDR + 11 * Lo.G;
  v0 += (64 - (((uintptr_t)v0 / sizeof(double)) & 63)) & 63;  // 512-byte aligned vectors
  P.r[0] = v0;
  P.r[1] = v0 + Lo.nvec;
  P.w[0] = v0 + 2 * Lo.nvec;
  P.w[1] = v0 + 3 * Lo.nvec;
  P.s[0] = v0 + 4 * Lo.nvec;
  P.s[1] = v0 + 5 * Lo.nvec;
  P.p = v0 + 6 * Lo.nvec;
  double* x_perm = v0 + 7 * Lo.nvec;
  double hst[SS_SIZE] = {0};
  const int bs = block_size_for(n);
  const int32_t nref = mx ? mx->refinements : 0;
  const double rin = mx ? mx->rtol_inner : rtol;
  int rc;
#define SMX_PCG_CALL(BS, WM, EX, V, AV, CP, IV)                                                   \
  run_pcg<BS, WM, EX, V, AV, CP>(n, nd, width, ld, cols, vals, IV, dinv, mode, b, b_part, nb_part,\
                                 project, raw_width, raw_ld, raw_cols, raw_vals, bpos, gb, x, warm, \
                                 perm, x_perm, rtol, atol, max_it, check_every, expected_its,     \
                                 nref, rin, P, st, hst)
#define SMX_PCG_WIDTH(BS, V, AV, CP, IV)                                                          \
  (w7 ? SMX_PCG_CALL(BS, 7, true, V, AV, CP, IV)                                                  \
      : wmax <= 8 ? SMX_PCG_CALL(BS, 8, false, V, AV, CP, IV)                                     \
                  : wmax <= 16 ? SMX_PCG_CALL(BS, 16, false, V, AV, CP, IV)                       \
                               : SMX_PCG_CALL(BS, 32, false, V, AV, CP, IV))
  const int32_t wmax = (mode == 1 && raw_width > width) ? raw_width : width;
  const bool w7 = width == 7 && wmax == 7;  // the tetrahedral Kuhn-grid stencil
  if (mx) {
    const float* v32 = mx->vals32;
    rc = bs == 256 ? SMX_PCG_WIDTH(256, float, float, true, v32)
                   : SMX_PCG_WIDTH(512, float, float, true, v32);
  } else {
    rc = bs == 256 ? SMX_PCG_WIDTH(256, double, double, false, vals)
                   : SMX_PCG_WIDTH(512, double, double, false, vals);
  }
#undef SMX_PCG_WIDTH
#undef SMX_PCG_CALL
  if (rc != SMX_OK) return rc;
  if (check_every < 0) {  // speculative: nothing waited for, the status stays on the device
    if (its_out) *its_out = -1;
    return SMX_OK;
  }
  if (its_out) *its_out = (int32_t)(hst[SS_ITERS] + hst[SS_ITERS0]);
  if (rnorm_out) *rnorm_out = sqrt(hst[SS_RR]);
  if (sums_out) sums_out[0] = hst[SS_SUMX];
  if (hst[SS_FAIL] == 3.0) {
    smx::set_error("CG operator is not in solver form: some row does not hold its diagonal in "
                   "slot 0 (build it with smx_pcg_prepare)");
    return SMX_ERR_VALUE;
  }
  if (hst[SS_FAIL] == 1.0) {
    smx::set_error("CG breakdown after %d iterations (rnorm %g)", (int)hst[SS_ITERS],
                   sqrt(hst[SS_RR]));
    return SMX_ERR_NOT_CONVERGED;
  }
  if (hst[SS_FAIL] == 2.0 || hst[SS_DONE] == 0.0) {
    smx::set_error("CG did not converge in %d iterations (rnorm %g, tol %g)", max_it,
                   sqrt(hst[SS_RR]), sqrt(hst[SS_TOL2]));
    return SMX_ERR_NOT_CONVERGED;
  }
  return SMX_OK;
}

}  // namespace

extern "C" int smx_pcg_jacobi(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols,
                              const double* vals, const double* dinv, int32_t mode,
                              const double* b, const double* b_part, int64_t nb_part,
                              int32_t project, int32_t raw_width, int64_t raw_ld,
                              const int32_t* raw_cols, const double* raw_vals,
                              const int32_t* bpos, const double* gb, double* x, int32_t warm,
                              const int32_t* perm, double rtol, double atol, int32_t max_it,
                              int32_t check_every,
                              int32_t expected_its, double* work, int32_t* its_out,
                              double* rnorm_out, double* sums_out, void* stream) {
  return pcg_entry(n, nd, width, ld, cols, vals, dinv, mode, b, b_part, nb_part, project, raw_width,
                   raw_ld, raw_cols, raw_vals, bpos, gb, x, warm, perm, rtol, atol, max_it,
                   check_every, expected_its, work, its_out, rnorm_out, sums_out, nullptr, stream);
}

extern "C" int smx_pcg_jacobi_mixed(int64_t n, int64_t nd, int32_t width, int64_t ld, const int32_t* cols,
                                    const double* vals, const double* dinv, int32_t mode,
                                    const double* b, const double* b_part, int64_t nb_part,
                                    int32_t project, int32_t raw_width, int64_t raw_ld,
                                    const int32_t* raw_cols, const double* raw_vals,
                                    const int32_t* bpos, const double* gb, double* x,
                                    int32_t warm, const int32_t* perm, double rtol, double atol,
                                    int32_t max_it, int32_t check_every, int32_t expected_its,
                                    double* work, int32_t* its_out, double* rnorm_out,
                                    double* sums_out, const smx_cg_mixed* mixed, void* stream) {
  SMX_REQUIRE(mixed, SMX_ERR_VALUE, "null mixed-precision options");
  return pcg_entry(n, nd, width, ld, cols, vals, dinv, mode, b, b_part, nb_part, project, raw_width,
                   raw_ld, raw_cols, raw_vals, bpos, gb, x, warm, perm, rtol, atol, max_it,
                   check_every, expected_its, work, its_out, rnorm_out, sums_out, mixed, stream);
}
