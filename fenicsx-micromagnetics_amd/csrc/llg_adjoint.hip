// Per-node LLG right-hand side with per-cell Ms, and its vector-Jacobian
// product (the building block of the adjoint through the RK steps).
//
// Model (exchange + demag + uniform H0; m on nodes, Ms per cell):
//   Ms_i  = sum_{c ni i} Ms_c |c|/4 / V_i          (volume-weighted nodal Ms)
//   H_i   = cex_i (K m)_i + Hd_i + H0,  cex_i = -2A 1e18 / (mu0 Ms_i V_i)
//   r_i   = P (p m x H + alpha m x (m x H)) + S_i (1 - |m|^2) m,  S_i = s0 Ms_i
// with P = -gamma/(1+alpha^2) and s0 = gamma/(2(1+alpha^2)), as the reference's
// _LLG_RHS_KERNEL (llg_module_GPU.py:100-122) with Ms per node.  The field H
// and the demag part Hd come from smx_field_llg_rhs (per-row scaled exchange
// values, the demag gradient operator); these kernels add the per-node parts.
#include "smx_common.h"

namespace {

constexpr int kB = 256;

__device__ __forceinline__ void cross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

__global__ __launch_bounds__(kB) void node_rhs_kernel(int64_t n, const double* __restrict__ m,
                                                       const double* __restrict__ h,
                                                       const double* __restrict__ S, double P,
                                                       double alpha, double p, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const double mi[3] = {m[3 * i], m[3 * i + 1], m[3 * i + 2]};
  const double hi[3] = {h[3 * i], h[3 * i + 1], h[3 * i + 2]};
  double a[3], b[3];
  cross(mi, hi, a);
  cross(mi, a, b);
  const double st = S[i] * (1.0 - (mi[0] * mi[0] + mi[1] * mi[1] + mi[2] * mi[2]));
#pragma unroll
  for (int q = 0; q < 3; ++q) r[3 * i + q] = P * (p * a[q] + alpha * b[q]) + st * mi[q];
}

// Given lambda (the cotangent of r), with a = m x H:
//   g_a = P (p lambda + alpha lambda x m)
//   gH  = g_a x m                                      (d/dH)
//   gm  = H x g_a + P alpha (a x lambda) + S(1-|m|^2) lambda - 2 S (m.lambda) m
//   gk  = cex gH            (to be multiplied by K: the exchange transpose)
//   gd  = gH / V            (to be multiplied by G^T: the demag transpose)
//   gnode = gH . (-Hex / Ms) + s0 (1 - |m|^2)(m . lambda)   (d/dMs_i)
// where Hex = H - Hd - H0 = cex (K m).
__global__ __launch_bounds__(kB) void node_rhs_vjp_kernel(
    int64_t n, const double* __restrict__ m, const double* __restrict__ h,
    const double* __restrict__ hd, double h0x, double h0y, double h0z,
    const double* __restrict__ Ms, const double* __restrict__ S, double s0,
    const double* __restrict__ cex, const double* __restrict__ inv_vol, double P, double alpha,
    double p, const double* __restrict__ lam, double* __restrict__ gm, double* __restrict__ gk,
    double* __restrict__ gd, double* __restrict__ gnode) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const double mi[3] = {m[3 * i], m[3 * i + 1], m[3 * i + 2]};
  const double hi[3] = {h[3 * i], h[3 * i + 1], h[3 * i + 2]};
  const double li[3] = {lam[3 * i], lam[3 * i + 1], lam[3 * i + 2]};
  double a[3], lxm[3], ga[3], gH[3], hxg[3], axl[3];
  cross(mi, hi, a);
  cross(li, mi, lxm);
#pragma unroll
  for (int q = 0; q < 3; ++q) ga[q] = P * (p * li[q] + alpha * lxm[q]);
  cross(ga, mi, gH);
  cross(hi, ga, hxg);
  cross(a, li, axl);
  const double mm = mi[0] * mi[0] + mi[1] * mi[1] + mi[2] * mi[2];
  const double ml = mi[0] * li[0] + mi[1] * li[1] + mi[2] * li[2];
  const double Si = S[i];
#pragma unroll
  for (int q = 0; q < 3; ++q)
    gm[3 * i + q] = hxg[q] + P * alpha * axl[q] + Si * (1.0 - mm) * li[q] - 2.0 * Si * ml * mi[q];
  const double c = cex[i], iv = inv_vol[i];
  const double h0[3] = {h0x, h0y, h0z};
  double dex = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    gk[3 * i + q] = c * gH[q];
    gd[3 * i + q] = gH[q] * iv;
    dex += gH[q] * (hi[q] - hd[3 * i + q] - h0[q]);
  }
  gnode[i] = -dex / Ms[i] + s0 * (1.0 - mm) * ml;
}

// gcell[c] (+)= sum over the cell's 4 nodes of |c|/4 / V_i * gnode[i]
__global__ __launch_bounds__(kB) void node_to_cell_kernel(int64_t T, const int32_t* __restrict__ cells,
                                                          const double* __restrict__ qvol,
                                                          const double* __restrict__ inv_vol,
                                                          const double* __restrict__ gnode,
                                                          int32_t accumulate, double* __restrict__ gcell) {
  const int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (c >= T) return;
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int32_t i = cells[4 * c + a];
    s += gnode[i] * inv_vol[i];
  }
  s *= qvol[c];
  gcell[c] = accumulate ? gcell[c] + s : s;
}

// Ms_i = sum_{c ni i} |c|/4 Ms_c / V_i (incidence list in (node, cell) order)
__global__ __launch_bounds__(kB) void cell_to_node_kernel(int64_t N, const int64_t* __restrict__ inc_ptr,
                                                          const int32_t* __restrict__ inc,
                                                          const double* __restrict__ qvol,
                                                          const double* __restrict__ inv_vol,
                                                          const double* __restrict__ Ms_cell,
                                                          double* __restrict__ Ms_node) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int64_t e = inc_ptr[i]; e < inc_ptr[i + 1]; ++e) {
    const int32_t c = inc[e] >> 2;  // incidence entries are 4*cell + local vertex
    s += qvol[c] * Ms_cell[c];
  }
  Ms_node[i] = s * inv_vol[i];
}

// out[e] = vals[e] * scale[rows[e]] (rows[e] < 0: padding slot, out = 0)
__global__ __launch_bounds__(kB) void row_scale_kernel(int64_t ns, const double* __restrict__ vals,
                                                       const int32_t* __restrict__ rows,
                                                       const double* __restrict__ scale,
                                                       double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= ns) return;
  const int32_t r = rows[e];
  out[e] = r >= 0 ? vals[e] * scale[r] : 0.0;
}

unsigned grid(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

}  // namespace

extern "C" int smx_node_rhs(int64_t n, const double* m, const double* h, const double* S, double P,
                            double alpha, double do_precess, double* r, void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative size");
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(m && h && S && r, SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(node_rhs_kernel, dim3(grid(n)), dim3(kB), 0, smx::as_stream(stream), n, m, h, S, P,
                     alpha, do_precess, r);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_node_rhs_vjp(int64_t n, const double* m, const double* h, const double* hd,
                                const double* h0_u, const double* Ms, const double* S, double s0,
                                const double* cex, const double* inv_vol, double P, double alpha,
                                double do_precess, const double* lam, double* gm, double* gk,
                                double* gd, double* gnode, void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative size");
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(m && h && hd && h0_u && Ms && S && cex && inv_vol && lam && gm && gk && gd && gnode,
              SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(node_rhs_vjp_kernel, dim3(grid(n)), dim3(kB), 0, smx::as_stream(stream), n, m, h, hd,
                     h0_u[0], h0_u[1], h0_u[2], Ms, S, s0, cex, inv_vol, P, alpha, do_precess, lam, gm,
                     gk, gd, gnode);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_node_to_cell(int64_t T, const int32_t* cells, const double* qvol, const double* inv_vol,
                                const double* gnode, int32_t accumulate, double* gcell, void* stream) {
  SMX_REQUIRE(T >= 0, SMX_ERR_SHAPE, "negative size");
  if (T == 0) return SMX_OK;
  SMX_REQUIRE(cells && qvol && inv_vol && gnode && gcell, SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(node_to_cell_kernel, dim3(grid(T)), dim3(kB), 0, smx::as_stream(stream), T, cells,
                     qvol, inv_vol, gnode, accumulate, gcell);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_cell_to_node(int64_t N, const int64_t* inc_ptr, const int32_t* inc, const double* qvol,
                                const double* inv_vol, const double* Ms_cell, double* Ms_node, void* stream) {
  SMX_REQUIRE(N >= 0, SMX_ERR_SHAPE, "negative size");
  if (N == 0) return SMX_OK;
  SMX_REQUIRE(inc_ptr && inc && qvol && inv_vol && Ms_cell && Ms_node, SMX_ERR_VALUE,
              "null pointer argument");
  hipLaunchKernelGGL(cell_to_node_kernel, dim3(grid(N)), dim3(kB), 0, smx::as_stream(stream), N, inc_ptr,
                     inc, qvol, inv_vol, Ms_cell, Ms_node);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_row_scale(int64_t nslots, const double* vals, const int32_t* rows, const double* scale,
                             double* out, void* stream) {
  SMX_REQUIRE(nslots >= 0, SMX_ERR_SHAPE, "negative size");
  if (nslots == 0) return SMX_OK;
  SMX_REQUIRE(vals && rows && scale && out, SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(row_scale_kernel, dim3(grid(nslots)), dim3(kB), 0, smx::as_stream(stream), nslots, vals,
                     rows, scale, out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}
