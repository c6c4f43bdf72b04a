// Jacobi-preconditioned conjugate gradients for the two FEM Poisson solves.
//
// Chronopoulos-Gear formulation: both inner products of an iteration come
// from one reduction, so an iteration is ONE kernel launch:
//   s_i = w_i + b_i s_{i-1};  p_i = u_i + b_i p_{i-1};  x += a_i p_i
//   r_{i+1} = r_i - a_i s_i;  u = D^-1 r;  w_{i+1} = A u_{i+1}
//   g = (r,u), d = (w,u), rr = (r,r);  b = g/g_old;  a = g/(d - b g/a_old)
// The SpMV needs u_{i+1} at neighbour rows, which each thread rebuilds from
// the previous iterate (r, w, s are double-buffered), so no grid-wide barrier
// is needed inside the iteration.  The four sums (g, d, rr, sum x) are formed
// by the "last block" reduction in a fixed order, and that block also updates
// the scalars and the convergence flag in device memory: the next launch
// reads them, and once the flag is set every later launch is a no-op.  The
// host only synchronizes every `check_every` iterations.
#include "smx_common.h"

namespace {

constexpr int kBlock = 256;

enum {
  ST_ALPHA = 0,
  ST_BETA = 1,
  ST_GAMMA = 2,
  ST_RR = 3,
  ST_TOL2 = 4,
  ST_DONE = 5,
  ST_ITERS = 6,
  ST_BB = 7,
  ST_SUMX = 8,
  ST_FAIL = 9,
  ST_SIZE = 32
};

struct CGPtrs {
  double* st;
  double* red;
  double* r[2];
  double* w[2];
  double* s[2];
  double* p;
};

__device__ __forceinline__ void row_span(int64_t row, const int64_t* sp, int64_t& start,
                                         int64_t& w, int& lane) {
  const int64_t s = row >> 6;
  lane = (int)(row & 63);
  start = sp[s];
  w = (sp[s + 1] - start) >> 6;
}

__global__ void cg_reset_kernel(double* st) {
  st[ST_DONE] = 0.0;
  st[ST_ITERS] = 0.0;
  st[ST_FAIL] = 0.0;
}

// r0 = b - shift - A x0   (mode 0) ;  Dirichlet lifting (mode 1)
__global__ __launch_bounds__(kBlock) void cg_init_a_kernel(
    int64_t n, int32_t mode, const int64_t* __restrict__ sp, const int32_t* __restrict__ cols,
    const double* __restrict__ vals, const double* __restrict__ b,
    const double* __restrict__ b_sum, int32_t project, const int32_t* __restrict__ bpos,
    const double* __restrict__ gb, double* __restrict__ x, int32_t warm, CGPtrs P, double rtol,
    double atol) {
  __shared__ double scratch[kBlock / 64];
  const int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double bb = 0.0;
  if (row < n) {
    int64_t start, w;
    int lane;
    row_span(row, sp, start, w, lane);
    double r0, bk;
    if (mode == 0) {
      bk = b[row] - ((project && b_sum) ? (*b_sum) / (double)n : 0.0);
      double ax = 0.0;
      if (warm) {
        for (int64_t j = 0; j < w; ++j) {
          const int64_t e = start + j * 64 + lane;
          ax = fma(vals[e], x[cols[e]], ax);
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
        for (int64_t j = 0; j < w; ++j) {
          const int64_t e = start + j * 64 + lane;
          const int64_t c = cols[e];
          const double a = vals[e];
          const int32_t cp = bpos[c];
          if (cp >= 0) {
            const double gv = gb[cp];
            sx = fma(a, gv, sx);
            sg = fma(a, gv, sg);
          } else if (warm) {
            sx = fma(a, x[c], sx);
          }
        }
        if (!warm) x[row] = 0.0;
        r0 = -sx;
        bk = -sg;
      }
    }
    P.r[0][row] = r0;
    P.s[0][row] = 0.0;
    P.p[row] = 0.0;
    bb = bk * bk;
  }
  double v[1] = {bb};
  if (smx::grid_sum_last<1>(v, P.red, scratch) && threadIdx.x == 0) {
    P.st[ST_BB] = v[0];
    const double t1 = rtol * rtol * v[0];
    const double t2 = atol * atol;
    P.st[ST_TOL2] = t1 > t2 ? t1 : t2;
  }
}

// w0 = A D^-1 r0 and the first scalars
__global__ __launch_bounds__(kBlock) void cg_init_b_kernel(int64_t n,
                                                           const int64_t* __restrict__ sp,
                                                           const int32_t* __restrict__ cols,
                                                           const double* __restrict__ vals,
                                                           const double* __restrict__ dinv,
                                                           const double* __restrict__ x,
                                                           CGPtrs P) {
  __shared__ double scratch[4 * kBlock / 64];
  const int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (row < n) {
    int64_t start, w;
    int lane;
    row_span(row, sp, start, w, lane);
    const double* r = P.r[0];
    double acc = 0.0;
    for (int64_t j = 0; j < w; ++j) {
      const int64_t e = start + j * 64 + lane;
      const int64_t c = cols[e];
      acc = fma(vals[e], r[c] * dinv[c], acc);
    }
    P.w[0][row] = acc;
    const double rk = r[row];
    const double uk = rk * dinv[row];
    v[0] = rk * uk;
    v[1] = acc * uk;
    v[2] = rk * rk;
    v[3] = x[row];
  }
  if (smx::grid_sum_last<4>(v, P.red, scratch) && threadIdx.x == 0) {
    const double g = v[0], d = v[1];
    P.st[ST_ALPHA] = (d != 0.0) ? g / d : 0.0;
    P.st[ST_BETA] = 0.0;
    P.st[ST_GAMMA] = g;
    P.st[ST_RR] = v[2];
    P.st[ST_SUMX] = v[3];
    if (v[2] <= P.st[ST_TOL2] || g == 0.0) P.st[ST_DONE] = 1.0;
    if (!(d > 0.0) && !(v[2] <= P.st[ST_TOL2])) {
      P.st[ST_DONE] = 1.0;
      P.st[ST_FAIL] = 1.0;
    }
  }
}

__global__ __launch_bounds__(kBlock) void cg_iter_kernel(int64_t n,
                                                         const int64_t* __restrict__ sp,
                                                         const int32_t* __restrict__ cols,
                                                         const double* __restrict__ vals,
                                                         const double* __restrict__ dinv,
                                                         double* __restrict__ x, CGPtrs P,
                                                         int parity) {
  __shared__ double scratch[4 * kBlock / 64];
  if (P.st[ST_DONE] != 0.0) return;  // uniform over the grid
  const double alpha = P.st[ST_ALPHA];
  const double beta = P.st[ST_BETA];
  const int q = parity & 1;
  const double* __restrict__ rq = P.r[q];
  const double* __restrict__ wq = P.w[q];
  const double* __restrict__ sq = P.s[q];
  const int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (row < n) {
    int64_t start, w;
    int lane;
    row_span(row, sp, start, w, lane);
    const double dk = dinv[row];
    const double s_n = wq[row] + beta * sq[row];
    const double p_n = rq[row] * dk + beta * P.p[row];
    const double x_n = x[row] + alpha * p_n;
    const double r_n = rq[row] - alpha * s_n;
    const double u_n = r_n * dk;
    double acc = 0.0;
    for (int64_t j = 0; j < w; ++j) {
      const int64_t e = start + j * 64 + lane;
      const int64_t c = cols[e];
      const double rc = rq[c] - alpha * (wq[c] + beta * sq[c]);
      acc = fma(vals[e], rc * dinv[c], acc);
    }
    P.s[1 - q][row] = s_n;
    P.r[1 - q][row] = r_n;
    P.w[1 - q][row] = acc;
    P.p[row] = p_n;
    x[row] = x_n;
    v[0] = r_n * u_n;
    v[1] = acc * u_n;
    v[2] = r_n * r_n;
    v[3] = x_n;
  }
  if (smx::grid_sum_last<4>(v, P.red, scratch) && threadIdx.x == 0) {
    const double g_new = v[0], d_new = v[1];
    const double g_old = P.st[ST_GAMMA];
    const double b_new = g_new / g_old;
    const double den = d_new - b_new * g_new / alpha;
    const double a_new = g_new / den;
    P.st[ST_ALPHA] = a_new;
    P.st[ST_BETA] = b_new;
    P.st[ST_GAMMA] = g_new;
    P.st[ST_RR] = v[2];
    P.st[ST_SUMX] = v[3];
    P.st[ST_ITERS] = P.st[ST_ITERS] + 1.0;
    if (v[2] <= P.st[ST_TOL2]) {
      P.st[ST_DONE] = 1.0;
    } else if (!(g_new > 0.0) || !(den != 0.0) || !isfinite(a_new)) {
      P.st[ST_DONE] = 1.0;
      P.st[ST_FAIL] = 1.0;
    }
  }
}

int64_t pcg_grid(int64_t n) {
  const int64_t nslice = (n + 63) / 64;
  return (nslice * 64 + kBlock - 1) / kBlock;
}

}  // namespace

extern "C" int64_t smx_pcg_ws_size(int64_t n) {
  const int64_t G = pcg_grid(n) + 1;
  const int64_t red = smx::red_ws_doubles(G, 4);
  const int64_t nvec = (n + 31) / 32 * 32;
  return ST_SIZE + (red + 31) / 32 * 32 + 7 * nvec;
}

extern "C" int smx_pcg_jacobi(int64_t n, const int64_t* slice_ptr, const int32_t* cols,
                              const double* vals, const double* dinv, int32_t mode,
                              const double* b, const double* b_sum, int32_t project,
                              const int64_t* a_raw_slice_ptr, const int32_t* a_raw_cols,
                              const double* a_raw_vals, const int32_t* bpos, const double* gb,
                              double* x, int32_t warm, double rtol, double atol,
                              int32_t max_it, int32_t check_every, int32_t expected_its,
                              double* work, int32_t* its_out, double* rnorm_out,
                              double* sums_out, void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative size");
  SMX_REQUIRE(mode == 0 || mode == 1, SMX_ERR_VALUE, "mode must be 0 or 1");
  SMX_REQUIRE(max_it >= 0, SMX_ERR_VALUE, "max_it must be >= 0");
  if (its_out) *its_out = 0;
  if (rnorm_out) *rnorm_out = 0.0;
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(slice_ptr && cols && vals && dinv && x && work, SMX_ERR_VALUE,
              "null pointer argument");
  SMX_REQUIRE(mode == 1 || b, SMX_ERR_VALUE, "mode 0 needs b");
  SMX_REQUIRE(mode == 0 || (a_raw_slice_ptr && a_raw_cols && a_raw_vals && bpos && gb),
              SMX_ERR_VALUE, "mode 1 needs the raw operator, bpos and gb");
  hipStream_t st = smx::as_stream(stream);
  const int64_t G = pcg_grid(n);
  const int64_t red = smx::red_ws_doubles(G + 1, 4);
  const int64_t nvec = (n + 31) / 32 * 32;
  CGPtrs P;
  P.st = work;
  P.red = work + ST_SIZE;
  double* v0 = work + ST_SIZE + (red + 31) / 32 * 32;
  P.r[0] = v0;
  P.r[1] = v0 + nvec;
  P.w[0] = v0 + 2 * nvec;
  P.w[1] = v0 + 3 * nvec;
  P.s[0] = v0 + 4 * nvec;
  P.s[1] = v0 + 5 * nvec;
  P.p = v0 + 6 * nvec;

  hipLaunchKernelGGL(cg_reset_kernel, dim3(1), dim3(1), 0, st, P.st);
  SMX_LAUNCH_CHECK();
  if (mode == 0) {
    hipLaunchKernelGGL(cg_init_a_kernel, dim3(G), dim3(kBlock), 0, st, n, mode, slice_ptr, cols,
                       vals, b, b_sum, project, bpos, gb, x, warm, P, rtol, atol);
  } else {
    hipLaunchKernelGGL(cg_init_a_kernel, dim3(G), dim3(kBlock), 0, st, n, mode, a_raw_slice_ptr,
                       a_raw_cols, a_raw_vals, b, b_sum, project, bpos, gb, x, warm, P, rtol,
                       atol);
  }
  SMX_LAUNCH_CHECK();
  hipLaunchKernelGGL(cg_init_b_kernel, dim3(G), dim3(kBlock), 0, st, n, slice_ptr, cols, vals,
                     dinv, x, P);
  SMX_LAUNCH_CHECK();

  if (check_every < 1) check_every = 16;
  int32_t batch = expected_its > 0 ? expected_its : check_every;
  int32_t launched = 0;
  double hst[ST_SIZE];
  bool done = false;
  SMX_HIP_CHECK(hipMemcpyAsync(hst, P.st, sizeof(double) * 10, hipMemcpyDeviceToHost, st));
  SMX_HIP_CHECK(hipStreamSynchronize(st));
  done = hst[ST_DONE] != 0.0;
  while (!done && launched < max_it) {
    int32_t nb = batch < (max_it - launched) ? batch : (max_it - launched);
    for (int32_t t = 0; t < nb; ++t) {
      hipLaunchKernelGGL(cg_iter_kernel, dim3(G), dim3(kBlock), 0, st, n, slice_ptr, cols, vals,
                         dinv, x, P, (launched + t) & 1);
    }
    SMX_LAUNCH_CHECK();
    launched += nb;
    SMX_HIP_CHECK(hipMemcpyAsync(hst, P.st, sizeof(double) * 10, hipMemcpyDeviceToHost, st));
    SMX_HIP_CHECK(hipStreamSynchronize(st));
    done = hst[ST_DONE] != 0.0;
    batch = check_every;
  }
  if (its_out) *its_out = (int32_t)hst[ST_ITERS];
  if (rnorm_out) *rnorm_out = sqrt(hst[ST_RR]);
  if (sums_out) sums_out[0] = hst[ST_SUMX];
  if (hst[ST_FAIL] != 0.0) {
    smx::set_error("CG breakdown after %d iterations (rnorm %g)", (int)hst[ST_ITERS],
                   sqrt(hst[ST_RR]));
    return SMX_ERR_NOT_CONVERGED;
  }
  if (!done) {
    smx::set_error("CG did not converge in %d iterations (rnorm %g, tol %g)", max_it,
                   sqrt(hst[ST_RR]), sqrt(hst[ST_TOL2]));
    return SMX_ERR_NOT_CONVERGED;
  }
  return SMX_OK;
}
