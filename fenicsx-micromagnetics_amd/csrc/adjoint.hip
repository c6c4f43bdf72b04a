// Energy of the micromagnetic state with per-cell Ms and A, and its adjoint.
//
//   E_ex = 1e-9  sum_c A_c  sum_ab K^c_ab m_a.m_b          (K^c element stiffness)
//   E_an = -1e-27 Ku sum_i V_i (m_i.n_i)^2
//   E_d  = 1/2 mu0 1e-27 b^T phi,  b = Div_Ms m,  phi = L b
// (for uniform Ms, E_d equals the lumped -1/2 mu0 Ms sum V_i m_i.H_i of the
// reference, since G = -Div^T / Ms).  Cell kernels evaluate per-cell sums; node
// kernels gather them over the node -> (cell, local vertex) incidence list, so
// nothing is accumulated with atomics and every result is deterministic.
//
// Per-cell data: cells (T x 4), cgv = vol/4 grad(lambda_a) (T x 4 x 3),
// ck = vol grad(lambda_a).grad(lambda_b) (T x 16).
#include "smx_common.h"

namespace {

constexpr int kBlock = 256;

// s_c = Ms_c sum_b m_b;  f_{c,a} = 2e-9 A_c sum_b K_ab m_b;  e_c = 1e-9 A_c m^T K m
__global__ __launch_bounds__(kBlock) void cell_forward_kernel(
    int64_t T, const int32_t* __restrict__ cells, const double* __restrict__ ck,
    const double* __restrict__ Ms, const double* __restrict__ A, const double* __restrict__ m,
    double* __restrict__ s, double* __restrict__ f, double* __restrict__ e,
    double* __restrict__ partials) {
  __shared__ double scr[kBlock / 64];
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double ec = 0.0;
  if (c < T) {
    double mv[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t v = cells[4 * c + a];
#pragma unroll
      for (int d = 0; d < 3; ++d) mv[a][d] = m[3 * v + d];
    }
    const double ms = Ms[c], ac = A ? A[c] : 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) s[3 * c + d] = ms * (mv[0][d] + mv[1][d] + mv[2][d] + mv[3][d]);
    double quad = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      double km[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const double k = ck[16 * c + 4 * a + b];
#pragma unroll
        for (int d = 0; d < 3; ++d) km[d] = fma(k, mv[b][d], km[d]);
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) f[12 * c + 3 * a + d] = 2e-9 * ac * km[d];
      quad += mv[a][0] * km[0] + mv[a][1] * km[1] + mv[a][2] * km[2];
    }
    ec = 1e-9 * ac * quad;
    e[c] = 1e-9 * quad;
  }
  if (partials) {
    double v[1] = {ec};
    smx::block_sum<1>(v, scr);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
  }
}

// b_a = sum_{(c,l) at a} cgv[c,l,:] . s_c  (divergence with per-cell Ms)
__global__ __launch_bounds__(kBlock) void node_div_kernel(int64_t N, const int64_t* __restrict__ ptr,
                                                          const int32_t* __restrict__ inc,
                                                          const double* __restrict__ cgv,
                                                          const double* __restrict__ s,
                                                          double* __restrict__ b,
                                                          double* __restrict__ partials) {
  __shared__ double scr[kBlock / 64];
  const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double acc = 0.0;
  if (a < N) {
    for (int64_t e = ptr[a]; e < ptr[a + 1]; ++e) {
      const int64_t cl = inc[e];
      const int64_t c = cl >> 2;
      const double* g = cgv + 3 * cl;
      acc += g[0] * s[3 * c] + g[1] * s[3 * c + 1] + g[2] * s[3 * c + 2];
    }
    b[a] = acc;
  }
  if (partials) {
    double v[1] = {acc};
    smx::block_sum<1>(v, scr);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
  }
}

// t_c = Ms_c sum_a y_a cgv[c,a,:];  dMs_c = (sum_a y_a cgv[c,a,:]) . sum_b m_b
__global__ __launch_bounds__(kBlock) void cell_backward_kernel(
    int64_t T, const int32_t* __restrict__ cells, const double* __restrict__ cgv,
    const double* __restrict__ Ms, const double* __restrict__ y, const double* __restrict__ m,
    double* __restrict__ t, double* __restrict__ dMs) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= T) return;
  double gy[3] = {0.0, 0.0, 0.0}, sm[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t v = cells[4 * c + a];
    const double ya = y[v];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      gy[d] = fma(ya, cgv[12 * c + 3 * a + d], gy[d]);
      sm[d] += m[3 * v + d];
    }
  }
  const double ms = Ms[c];
#pragma unroll
  for (int d = 0; d < 3; ++d) t[3 * c + d] = ms * gy[d];
  dMs[c] = gy[0] * sm[0] + gy[1] * sm[1] + gy[2] * sm[2];
}

// grad_a = sum_{(c,l) at a} (f[c,l] + t_c)  - 2e-27 Ku V_a (m_a.n_a) n_a ;  E_an partials
__global__ __launch_bounds__(kBlock) void node_grad_kernel(
    int64_t N, const int64_t* __restrict__ ptr, const int32_t* __restrict__ inc,
    const double* __restrict__ f, const double* __restrict__ t, const double* __restrict__ m,
    const double* __restrict__ vol, const double* __restrict__ axis, double ax0, double ax1,
    double ax2, double Ku, double* __restrict__ grad, double* __restrict__ partials) {
  __shared__ double scr[kBlock / 64];
  const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double ea = 0.0;
  if (a < N) {
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int64_t e = ptr[a]; e < ptr[a + 1]; ++e) {
      const int64_t cl = inc[e];
      const int64_t c = cl >> 2;
      if (f) {
        g0 += f[3 * cl];
        g1 += f[3 * cl + 1];
        g2 += f[3 * cl + 2];
      }
      if (t) {
        g0 += t[3 * c];
        g1 += t[3 * c + 1];
        g2 += t[3 * c + 2];
      }
    }
    if (Ku != 0.0) {
      const double nx = axis ? axis[3 * a] : ax0, ny = axis ? axis[3 * a + 1] : ax1,
                   nz = axis ? axis[3 * a + 2] : ax2;
      const double mn = m[3 * a] * nx + m[3 * a + 1] * ny + m[3 * a + 2] * nz;
      const double k = -2e-27 * Ku * vol[a] * mn;
      g0 += k * nx;
      g1 += k * ny;
      g2 += k * nz;
      ea = -1e-27 * Ku * vol[a] * mn * mn;
    }
    grad[3 * a] = g0;
    grad[3 * a + 1] = g1;
    grad[3 * a + 2] = g2;
  }
  if (partials) {
    double v[1] = {ea};
    smx::block_sum<1>(v, scr);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
  }
}

// Generic glue of the adjoint (one pass each):
//  op 0: out_i = bpos_i >= 0 ? 0 : x_i                          (interior restriction)
//  op 1: out_k = x[idx_k] - y[idx_k]   (k < n = Nb)               (boundary residual)
//  op 2: out_i = x_i + (bpos_i >= 0 ? y[bpos_i] : 0)             (boundary scatter-add)
//  op 3: out_i = a*(x_i - sx/n) + b*(y_i - sy/n)                 (projected sums)
__global__ __launch_bounds__(kBlock) void glue_kernel(int32_t op, int64_t n, const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const int32_t* __restrict__ idx, double a, double b,
                                                      const double* __restrict__ sx,
                                                      const double* __restrict__ sy, double nrm,
                                                      double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  switch (op) {
    case 0:
      out[i] = idx[i] >= 0 ? 0.0 : x[i];
      break;
    case 1: {
      const int64_t k = idx[i];
      out[i] = x[k] - y[k];
      break;
    }
    case 2: {
      const int32_t k = idx[i];
      out[i] = x[i] + (k >= 0 ? y[k] : 0.0);
      break;
    }
    default: {
      const double mx = sx ? (*sx) / nrm : 0.0;
      const double my = sy ? (*sy) / nrm : 0.0;
      out[i] = a * (x[i] - mx) + (y ? b * (y[i] - my) : 0.0);
      break;
    }
  }
}

__global__ __launch_bounds__(kBlock) void dot_partials_kernel(int64_t n, const double* __restrict__ x,
                                                              const double* __restrict__ y,
                                                              double* __restrict__ partials) {
  __shared__ double scr[kBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double v[1] = {i < n ? x[i] * y[i] : 0.0};
  smx::block_sum<1>(v, scr);
  if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

}  // namespace

extern "C" int64_t smx_energy_blocks(int64_t n) { return (n + kBlock - 1) / kBlock; }

extern "C" int smx_energy_cell_forward(int64_t T, const int32_t* cells, const double* ck,
                                       const double* Ms, const double* A, const double* m, double* s,
                                       double* f, double* e, double* partials, void* stream) {
  SMX_REQUIRE(T >= 0, SMX_ERR_SHAPE, "negative cell count");
  if (T == 0) return SMX_OK;
  SMX_REQUIRE(cells && ck && Ms && m && s && f && e, SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(cell_forward_kernel, dim3(smx::grid_for(T, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), T, cells, ck, Ms, A, m, s, f, e, partials);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_energy_node_div(int64_t N, const int64_t* ptr, const int32_t* inc, const double* cgv,
                                   const double* s, double* b, double* partials, void* stream) {
  SMX_REQUIRE(N >= 0, SMX_ERR_SHAPE, "negative node count");
  if (N == 0) return SMX_OK;
  SMX_REQUIRE(ptr && inc && cgv && s && b, SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(node_div_kernel, dim3(smx::grid_for(N, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), N, ptr, inc, cgv, s, b, partials);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_energy_cell_backward(int64_t T, const int32_t* cells, const double* cgv,
                                        const double* Ms, const double* y, const double* m, double* t,
                                        double* dMs, void* stream) {
  SMX_REQUIRE(T >= 0, SMX_ERR_SHAPE, "negative cell count");
  if (T == 0) return SMX_OK;
  SMX_REQUIRE(cells && cgv && Ms && y && m && t && dMs, SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(cell_backward_kernel, dim3(smx::grid_for(T, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), T, cells, cgv, Ms, y, m, t, dMs);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_energy_node_grad(int64_t N, const int64_t* ptr, const int32_t* inc, const double* f,
                                    const double* t, const double* m, const double* vol,
                                    const double* axis, const double* axis_u, double Ku, double* grad,
                                    double* partials, void* stream) {
  SMX_REQUIRE(N >= 0, SMX_ERR_SHAPE, "negative node count");
  if (N == 0) return SMX_OK;
  SMX_REQUIRE(ptr && inc && m && grad && (Ku == 0.0 || (vol && (axis || axis_u))), SMX_ERR_VALUE,
              "null pointer argument");
  const double a0 = axis_u ? axis_u[0] : 0.0, a1 = axis_u ? axis_u[1] : 0.0, a2 = axis_u ? axis_u[2] : 0.0;
  hipLaunchKernelGGL(node_grad_kernel, dim3(smx::grid_for(N, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), N, ptr, inc, f, t, m, vol, axis, a0, a1, a2, Ku, grad,
                     partials);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_vec_glue(int32_t op, int64_t n, const double* x, const double* y, const int32_t* idx,
                            double a, double b, const double* sx, const double* sy, double nrm,
                            double* out, void* stream) {
  SMX_REQUIRE(op >= 0 && op <= 3, SMX_ERR_VALUE, "bad op %d", op);
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative size");
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(x && out && (op == 3 || idx) && (op == 0 || op == 3 || y), SMX_ERR_VALUE,
              "null pointer argument");
  hipLaunchKernelGGL(glue_kernel, dim3(smx::grid_for(n, kBlock)), dim3(kBlock), 0, smx::as_stream(stream),
                     op, n, x, y, idx, a, b, sx, sy, nrm, out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_dot_partials(int64_t n, const double* x, const double* y, double* partials,
                                void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative size");
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(x && y && partials, SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(dot_partials_kernel, dim3(smx::grid_for(n, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), n, x, y, partials);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}
