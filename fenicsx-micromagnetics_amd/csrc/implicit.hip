// Implicit LLG time stepping: the 3x3 block-Jacobi preconditioner of the
// Newton-Krylov solves and the vector kernels of the restarted GMRES.
//
// The preconditioner is, node by node, the inverse of A_i = (shift + eps) I -
// J_i, J_i the block of JvContext.update_pc_full_fast_gpu and its kernel
// (llg_module_GPU.py:165-290, 1447-1505).  For the LLG right-hand side
//   r = -c1 m x H - c2 m x (m x H) + S (1 - |m|^2) m,
//   c1 = gamma p/(1+alpha^2), c2 = gamma alpha/(1+alpha^2),
// with the field's self-derivative taken isotropic, dH_i/dm_i = kappa I, the
// exact local derivatives are
//   d(m x H)/dm         = -[H]x + kappa [m]x          ([v]x u = v x u)
//   d(m x (m x H))/dm   = B + kappa (m m^T - |m|^2 I),
//       B_rr = m.H - H_r m_r,  B_rc = m_r H_c - 2 H_r m_c (r != c)
//   d(S (1-|m|^2) m)/dm = S ((1 - |m|^2) I - 2 m m^T).
// kappa is the mean (or mean absolute value) of the diagonal of the linear
// field operator (exchange, DMI and uniaxial self-terms plus the cubic
// anisotropy's diagonal derivative).  exact = 1 uses these derivatives with
// the RHS's signs (J = dr/dm); exact = 0 reproduces the reference's block,
// whose precession and damping parts carry the opposite sign
// (J = +c1 d(m x H)/dm + c2 d(m x (m x H))/dm + stabilization).  The 3x3
// inverse is by cofactors with |det| floored at det_eps.
#include "smx_common.h"

namespace {

constexpr int kBlock = 256;

// Diagonal of the linear field operator at node k, per component.
__device__ __forceinline__ void field_self_diag(const smx_field_args& a, int64_t k, double (&d)[3]) {
  d[0] = d[1] = d[2] = 0.0;
  if (a.ex_vals || a.dmi_vals) {
    const int64_t s = k >> 6;
    const int lane = (int)(k & 63);
    const int64_t start = a.slice_ptr[s];
    const int64_t w = (a.slice_ptr[s + 1] - start) >> 6;
    for (int64_t j = 0; j < w; ++j) {
      const int64_t e = start + j * 64 + lane;
      if (a.cols[e] != (int32_t)k) continue;
      if (a.ex_vals) {
        d[0] += a.ex_vals[e];
        d[1] += a.ex_vals[e];
        d[2] += a.ex_vals[e];
      }
      if (a.dmi_vals) {
        const int nc = a.dmi_ncomp > 0 ? a.dmi_ncomp : 9;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int o = a.dmi_ncomp > 0 ? a.dmi_off[4 * c] : 4 * c;  // diagonal component (c, c)
          if (o >= 0) d[c] += a.dmi_vals[(int64_t)nc * e + o];
        }
      }
    }
  }
  if (a.ani_pref != 0.0) {
    double n[3];
    for (int c = 0; c < 3; ++c) n[c] = a.ani_axis ? a.ani_axis[3 * k + c] : a.ani_axis_u[c];
    for (int c = 0; c < 3; ++c) d[c] += a.ani_pref * n[c] * n[c];
  }
  if (a.cub_pref != 0.0) {
    // diagonal of dH_cub/dm: component c of the derivative along e_c
    double cx[9];
    for (int i = 0; i < 9; ++i) cx[i] = a.cub_axes ? a.cub_axes[9 * k + i] : a.cub_axes_u[i];
    const double mx = a.m[3 * k], my = a.m[3 * k + 1], mz = a.m[3 * k + 2];
    const double a1 = mx * cx[0] + my * cx[1] + mz * cx[2];
    const double a2 = mx * cx[3] + my * cx[4] + mz * cx[5];
    const double a3 = mx * cx[6] + my * cx[7] + mz * cx[8];
    for (int c = 0; c < 3; ++c) {
      const double d1 = cx[c], d2 = cx[3 + c], d3 = cx[6 + c];
      const double s1 = d1 * (a2 * a2 + a3 * a3) + 2.0 * a1 * (a2 * d2 + a3 * d3);
      const double s2 = d2 * (a3 * a3 + a1 * a1) + 2.0 * a2 * (a3 * d3 + a1 * d1);
      const double s3 = d3 * (a1 * a1 + a2 * a2) + 2.0 * a3 * (a1 * d1 + a2 * d2);
      d[c] += a.cub_pref * (s1 * cx[c] + s2 * cx[3 + c] + s3 * cx[6 + c]);
    }
  }
}

__global__ __launch_bounds__(kBlock) void pc_build_kernel(smx_field_args a, const double* __restrict__ hm,
                                                          double shift, double eps_reg, double det_eps,
                                                          int32_t include_stab, int32_t use_abs_kappa,
                                                          int32_t exact, double* __restrict__ inv) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.n) return;
  double dg[3];
  field_self_diag(a, k, dg);
  const double kappa = use_abs_kappa ? (fabs(dg[0]) + fabs(dg[1]) + fabs(dg[2])) / 3.0
                                     : (dg[0] + dg[1] + dg[2]) / 3.0;
  const double m[3] = {a.m[3 * k], a.m[3 * k + 1], a.m[3 * k + 2]};
  const double h[3] = {hm[3 * k], hm[3 * k + 1], hm[3 * k + 2]};
  const double sgn = exact ? -1.0 : 1.0;
  const double c1 = -a.P * a.do_precess * sgn, c2 = -a.P * a.alpha * sgn;
  const double mh = m[0] * h[0] + m[1] * h[1] + m[2] * h[2];
  const double mm = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
  double J[3][3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) {
      // precession: +-c1 d(m x H)/dm = c1 (-[H]x + kappa [m]x), c1 carrying the sign
      double hx = 0.0, mx = 0.0;  // ([H]x)_rc and ([m]x)_rc, [v]x u = v x u
      const int t = 3 - r - c;    // the third index when r != c
      if (r != c) {
        const double sg = ((r + 1) % 3 == c) ? -1.0 : 1.0;  // [v]x_rc = -eps_rct v_t
        hx = sg * h[t];
        mx = sg * m[t];
      }
      const double prec = c1 * (-hx + kappa * mx);
      // damping: +-c2 d(m x (m x H))/dm = c2 (B + kappa C)
      const double B = (r == c) ? (mh - h[r] * m[c]) : (m[r] * h[c] - 2.0 * h[r] * m[c]);
      const double Cm = m[r] * m[c] - (r == c ? mm : 0.0);
      const double damp = c2 * (B + kappa * Cm);
      double stab = 0.0;
      if (include_stab) stab = a.stab * ((r == c ? 1.0 - mm : 0.0) - 2.0 * m[r] * m[c]);
      J[r][c] = prec + damp + stab;
    }
  }
  const double s = shift + eps_reg;
  double A[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = (r == c ? s : 0.0) - J[r][c];
  double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
               A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
               A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  if (fabs(det) < det_eps) det += (det + det_eps > 0.0 ? 1.0 : -1.0) * det_eps;
  const double id = 1.0 / det;
  double* o = inv + 9 * k;
  o[0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id;
  o[1] = -(A[0][1] * A[2][2] - A[0][2] * A[2][1]) * id;
  o[2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  o[3] = -(A[1][0] * A[2][2] - A[1][2] * A[2][0]) * id;
  o[4] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  o[5] = -(A[0][0] * A[1][2] - A[0][2] * A[1][0]) * id;
  o[6] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id;
  o[7] = -(A[0][0] * A[2][1] - A[0][1] * A[2][0]) * id;
  o[8] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
}

__global__ __launch_bounds__(kBlock) void block3_apply_kernel(int64_t n, const double* __restrict__ inv,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ y) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const double* o = inv + 9 * k;
  const double x0 = x[3 * k], x1 = x[3 * k + 1], x2 = x[3 * k + 2];
  y[3 * k] = o[0] * x0 + o[1] * x1 + o[2] * x2;
  y[3 * k + 1] = o[3] * x0 + o[4] * x1 + o[5] * x2;
  y[3 * k + 2] = o[6] * x0 + o[7] * x1 + o[8] * x2;
}

// part[j*G + b] = sum over workgroup b's entries of V_j . w (j < k), fixed
// order.  Each thread keeps kMdEpt entries of w in registers and sums kMdJ
// vectors at a time, so a workgroup needs one barrier per kMdJ vectors
// (double-buffered LDS) instead of two per vector.
constexpr int kMdEpt = 1;
constexpr int kMdJ = 32;
constexpr int kMdSpan = kBlock * kMdEpt;

__global__ __launch_bounds__(kBlock) void multi_dot_kernel(int64_t len, int32_t k, const double* __restrict__ V,
                                                           int64_t ldv, const double* __restrict__ w,
                                                           double* __restrict__ part) {
  __shared__ double scr[2][kMdJ][kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kMdSpan + threadIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double wv[kMdEpt];
#pragma unroll
  for (int q = 0; q < kMdEpt; ++q) {
    const int64_t i = base + (int64_t)q * kBlock;
    wv[q] = i < len ? w[i] : 0.0;
  }
  int buf = 0;
  for (int32_t j0 = 0; j0 < k; j0 += kMdJ, buf ^= 1) {
    double acc[kMdJ];
#pragma unroll
    for (int jj = 0; jj < kMdJ; ++jj) {
      acc[jj] = 0.0;
      if (j0 + jj < k) {
        const double* __restrict__ Vj = V + (int64_t)(j0 + jj) * ldv;
#pragma unroll
        for (int q = 0; q < kMdEpt; ++q) {
          const int64_t i = base + (int64_t)q * kBlock;
          if (i < len) acc[jj] = fma(Vj[i], wv[q], acc[jj]);
        }
      }
    }
#pragma unroll
    for (int jj = 0; jj < kMdJ; ++jj) {
      const double t = smx::wave_total(acc[jj]);
      if (lane == 0) scr[buf][jj][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < kMdJ && j0 + (int)threadIdx.x < k) {
      double sum = 0.0;
#pragma unroll
      for (int q = 0; q < kBlock / 64; ++q) sum += scr[buf][threadIdx.x][q];
      part[(int64_t)(j0 + threadIdx.x) * gridDim.x + blockIdx.x] = sum;
    }
    // the other LDS buffer is written next: no second barrier needed
  }
}

// out[j] = sum over b of part[j*G + b]: one workgroup per j, fixed order.
__global__ __launch_bounds__(kBlock) void multi_dot_reduce_kernel(int32_t k, int64_t G,
                                                                  const double* __restrict__ part,
                                                                  double* __restrict__ out) {
  __shared__ double scr[kBlock / 64];
  const int32_t j = blockIdx.x;
  double v = 0.0;
  for (int64_t b = threadIdx.x; b < G; b += kBlock) v += part[(int64_t)j * G + b];
  v = smx::wave_total(v);
  if ((threadIdx.x & 63) == 0) scr[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) s += scr[q];
    out[j] = s;
  }
}

// out = beta*out + sum_j c_j V_j  (c on the device)
__global__ __launch_bounds__(kBlock) void lincomb_kernel(int64_t len, int32_t k, const double* __restrict__ V,
                                                         int64_t ldv, const double* __restrict__ c,
                                                         double beta, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= len) return;
  double s = beta == 0.0 ? 0.0 : beta * out[i];
  for (int32_t j = 0; j < k; ++j) s = fma(c[j], V[j * ldv + i], s);
  out[i] = s;
}

}  // namespace

extern "C" int smx_llg_pc_build(const smx_field_args* a, const double* hm, double shift, double eps_reg,
                                double det_eps, int32_t include_stab, int32_t use_abs_kappa, int32_t exact,
                                double* inv, void* stream) {
  SMX_REQUIRE(a && hm && inv && a->m, SMX_ERR_VALUE, "null pointer argument");
  SMX_REQUIRE(a->n >= 0, SMX_ERR_SHAPE, "negative n");
  SMX_REQUIRE(!(a->ex_vals || a->dmi_vals) || (a->slice_ptr && a->cols), SMX_ERR_VALUE,
              "graph pattern missing");
  if (a->n == 0) return SMX_OK;
  hipLaunchKernelGGL(pc_build_kernel, dim3(smx::grid_for(a->n, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), *a, hm, shift, eps_reg, det_eps, include_stab, use_abs_kappa,
                     exact, inv);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_block3_apply(int64_t n, const double* inv, const double* x, double* y, void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative n");
  SMX_REQUIRE(inv && x && y, SMX_ERR_VALUE, "null pointer argument");
  if (n == 0) return SMX_OK;
  hipLaunchKernelGGL(block3_apply_kernel, dim3(smx::grid_for(n, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), n, inv, x, y);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int64_t smx_multi_dot_ws_size(int64_t len, int32_t k) {
  return (int64_t)smx::grid_for(len, kMdSpan) * (k > 0 ? k : 1);
}

extern "C" int smx_multi_dot(int64_t len, int32_t k, const double* V, int64_t ldv, const double* w,
                             double* out, double* work, void* stream) {
  SMX_REQUIRE(len >= 0 && k >= 0 && k <= 1024 && ldv >= len, SMX_ERR_SHAPE, "bad sizes");
  SMX_REQUIRE((k == 0 || (V && w)) && out && work, SMX_ERR_VALUE, "null pointer argument");
  if (k == 0) return SMX_OK;
  hipStream_t st = smx::as_stream(stream);
  if (len == 0) {
    hipLaunchKernelGGL(multi_dot_reduce_kernel, dim3((unsigned)k), dim3(kBlock), 0, st, k, (int64_t)0, work, out);
    SMX_LAUNCH_CHECK();
    return SMX_OK;
  }
  const int G = smx::grid_for(len, kMdSpan);
  hipLaunchKernelGGL(multi_dot_kernel, dim3(G), dim3(kBlock), 0, st, len, k, V, ldv, w, work);
  SMX_LAUNCH_CHECK();
  hipLaunchKernelGGL(multi_dot_reduce_kernel, dim3((unsigned)k), dim3(kBlock), 0, st, k, (int64_t)G, work, out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_lincomb(int64_t len, int32_t k, const double* V, int64_t ldv, const double* c, double beta,
                           double* out, void* stream) {
  SMX_REQUIRE(len >= 0 && k >= 0 && ldv >= len, SMX_ERR_SHAPE, "bad sizes");
  SMX_REQUIRE(out && (k == 0 || (V && c)), SMX_ERR_VALUE, "null pointer argument");
  if (len == 0) return SMX_OK;
  hipLaunchKernelGGL(lincomb_kernel, dim3(smx::grid_for(len, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), len, k, V, ldv, c, beta, out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}
