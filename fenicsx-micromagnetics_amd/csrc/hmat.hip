// H-matrix construction (batched full-pivot ACA) and matvec on gfx950.
//
// ACA: one 1024-thread workgroup per admissible block; the block stays in
// HBM/L2 and every rank step is ONE pass over it that applies the rank-1
// update, accumulates ||R||_F^2 and finds the next pivot (first argmax |R| in
// row-major order), so a block of rank r costs r+1 passes and no host
// round trip.  Waves own rows, lanes own columns (coalesced rows).
//
// Matvec: z_b = V_b xp[J_b] with one wavefront per far block, then one
// workgroup per leaf cluster produces its rows: diagonal jump + near CSR
// (one wavefront per row, lanes over entries) + sum_b U_b[rows,:] z_b (lanes
// are rows of the column-major U slice: 256 B coalesced per rank), reduced in
// LDS in a fixed order and scattered to the original boundary order.
#include <cfloat>

#include "smx_common.h"

namespace {

constexpr int kAcaThreads = 1024;
constexpr int kAcaWaves = kAcaThreads / 64;

__device__ __forceinline__ void am_merge(double& bv, int64_t& bi, double v, int64_t i) {
  if (v > bv || (v == bv && i < bi)) {
    bv = v;
    bi = i;
  }
}

// Block-wide (sum, argmax) reduction; results broadcast to all threads.
__device__ __forceinline__ void block_sum_argmax(double& s, double& bv, int64_t& bi,
                                                 double* ss, double* sv, int64_t* si) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off, 64);
    const double ov = __shfl_xor(bv, off, 64);
    const int64_t oi = __shfl_xor(bi, off, 64);
    am_merge(bv, bi, ov, oi);
  }
  if (lane == 0) {
    ss[wid] = s;
    sv[wid] = bv;
    si[wid] = bi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ts = 0.0, tv = -1.0;
    int64_t ti = INT64_MAX;
    for (int w = 0; w < kAcaWaves; ++w) {
      ts += ss[w];
      am_merge(tv, ti, sv[w], si[w]);
    }
    ss[kAcaWaves] = ts;
    sv[kAcaWaves] = tv;
    si[kAcaWaves] = ti;
  }
  __syncthreads();
  s = ss[kAcaWaves];
  bv = sv[kAcaWaves];
  bi = si[kAcaWaves];
  __syncthreads();
}

__global__ __launch_bounds__(kAcaThreads) void aca_kernel(double* __restrict__ A,
                                                          const smx_aca_task* __restrict__ tasks,
                                                          double rtol, double* __restrict__ U,
                                                          double* __restrict__ V,
                                                          int32_t* __restrict__ rank_out,
                                                          double* __restrict__ resid_out,
                                                          int32_t* __restrict__ status_out) {
  __shared__ double ss[kAcaWaves + 1];
  __shared__ double sv[kAcaWaves + 1];
  __shared__ int64_t si[kAcaWaves + 1];
  const smx_aca_task T = tasks[blockIdx.x];
  double* __restrict__ R = A + T.a_off;
  const int m = T.m, n = T.n;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;

  double s = 0.0, bv = -1.0;
  int64_t bi = INT64_MAX;
  for (int r = wid; r < m; r += kAcaWaves) {
    const double* row = R + (int64_t)r * n;
    for (int c = lane; c < n; c += 64) {
      const double v = row[c];
      s += v * v;
      am_merge(bv, bi, fabs(v), (int64_t)r * n + c);
    }
  }
  block_sum_argmax(s, bv, bi, ss, sv, si);
  const double normA = sqrt(s);
  int rank = 0, status = 2;
  double rel = 1.0;
  if (normA == 0.0 || T.rcap == 0) {
    status = 3;
    rel = 0.0;
  } else {
    for (int k = 0; k < T.rcap; ++k) {
      const int pi = (int)(bi / n);
      const int pj = (int)(bi - (int64_t)pi * n);
      const double piv = R[bi];
      if (fabs(piv) <= DBL_EPSILON * normA) {
        status = 1;
        break;
      }
      double* __restrict__ u = U + T.u_off + (int64_t)k * m;
      double* __restrict__ v = V + T.v_off + (int64_t)k * n;
      for (int r = threadIdx.x; r < m; r += kAcaThreads) u[r] = R[(int64_t)r * n + pj];
      for (int c = threadIdx.x; c < n; c += kAcaThreads) v[c] = R[(int64_t)pi * n + c] / piv;
      __threadfence_block();
      __syncthreads();
      s = 0.0;
      bv = -1.0;
      bi = INT64_MAX;
      for (int r = wid; r < m; r += kAcaWaves) {
        double* row = R + (int64_t)r * n;
        const double ur = u[r];
        for (int c = lane; c < n; c += 64) {
          const double nv = __dsub_rn(row[c], __dmul_rn(ur, v[c]));
          row[c] = nv;
          s += nv * nv;
          am_merge(bv, bi, fabs(nv), (int64_t)r * n + c);
        }
      }
      block_sum_argmax(s, bv, bi, ss, sv, si);
      rel = sqrt(s) / normA;
      rank = k + 1;
      if (rel <= rtol) {
        status = 0;
        break;
      }
    }
  }
  if (threadIdx.x == 0) {
    rank_out[blockIdx.x] = rank;
    resid_out[blockIdx.x] = rel;
    status_out[blockIdx.x] = status;
  }
}

__global__ void hm_gather_kernel(int64_t nb, const double* __restrict__ src,
                                 const int32_t* __restrict__ gidx, const double* __restrict__ shift,
                                 double scale, double* __restrict__ xp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const double sh = shift ? (*shift) * scale : 0.0;
  xp[i] = src[gidx[i]] - sh;
}

__device__ __forceinline__ double ld_factor(const void* p, int64_t i, int32_t dt) {
  return dt == SMX_F32 ? (double)reinterpret_cast<const float*>(p)[i]
                       : reinterpret_cast<const double*>(p)[i];
}

template <typename T>
__global__ __launch_bounds__(256) void hm_far_z_kernel(smx_hmatrix H) {
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= H.nfar) return;
  const int32_t c0 = H.far_c0[b], n = H.far_n[b], r = H.far_rank[b];
  const T* __restrict__ V = reinterpret_cast<const T*>(H.V) + H.far_v_off[b];
  const double* __restrict__ x = H.xp + c0;
  double* __restrict__ z = H.z + H.far_z_off[b];
  int k = 0;
  for (; k + 4 <= r; k += 4) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int j = lane; j < n; j += 64) {
      const double xj = x[j];
      s0 = fma((double)V[(int64_t)k * n + j], xj, s0);
      s1 = fma((double)V[(int64_t)(k + 1) * n + j], xj, s1);
      s2 = fma((double)V[(int64_t)(k + 2) * n + j], xj, s2);
      s3 = fma((double)V[(int64_t)(k + 3) * n + j], xj, s3);
    }
    s0 = smx::wave_sum(s0);
    s1 = smx::wave_sum(s1);
    s2 = smx::wave_sum(s2);
    s3 = smx::wave_sum(s3);
    if (lane == 0) {
      z[k] = s0;
      z[k + 1] = s1;
      z[k + 2] = s2;
      z[k + 3] = s3;
    }
  }
  for (; k < r; ++k) {
    double s0 = 0.0;
    for (int j = lane; j < n; j += 64) s0 = fma((double)V[(int64_t)k * n + j], x[j], s0);
    s0 = smx::wave_sum(s0);
    if (lane == 0) z[k] = s0;
  }
}

template <typename T, typename TN>
__global__ __launch_bounds__(256) void hm_leaf_kernel(smx_hmatrix H, double* __restrict__ y) {
  __shared__ double nacc[64];
  __shared__ double facc[4][64];
  const int64_t L = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int32_t l0 = H.leaf_start[L], l1 = H.leaf_stop[L];
  const int64_t e0 = H.leaf_fe_ptr[L], e1 = H.leaf_fe_ptr[L + 1];
  const T* __restrict__ U = reinterpret_cast<const T*>(H.U);
  const TN* __restrict__ NV = reinterpret_cast<const TN*>(H.near_val);
  for (int32_t rb = l0; rb < l1; rb += 64) {
    const int rows_here = (l1 - rb) < 64 ? (l1 - rb) : 64;
    for (int i = wid; i < rows_here; i += 4) {
      const int64_t row = rb + i;
      double s = 0.0;
      if (H.near_ptr) {
        for (int64_t e = H.near_ptr[row] + lane; e < H.near_ptr[row + 1]; e += 64)
          s = fma((double)NV[e], H.xp[H.near_idx[e]], s);
      }
      s = smx::wave_sum(s);
      if (lane == 0) nacc[i] = s;
    }
    double acc = 0.0;
    if (lane < rows_here) {
      for (int64_t e = e0 + wid; e < e1; e += 4) {
        const int32_t b = H.leaf_fe_block[e];
        const int64_t o = (int64_t)H.leaf_fe_row[e] + (rb - l0) + lane;
        const int32_t m = H.far_m[b], r = H.far_rank[b];
        const T* __restrict__ Ub = U + H.far_u_off[b] + o;
        const double* __restrict__ zb = H.z + H.far_z_off[b];
        for (int k = 0; k < r; ++k) acc = fma((double)Ub[(int64_t)k * m], zb[k], acc);
      }
    }
    facc[wid][lane] = acc;
    __syncthreads();
    if (threadIdx.x < rows_here) {
      const int i = threadIdx.x;
      const int64_t p = rb + i;
      const double v = H.diag_p[p] * H.xp[p] + nacc[i] + facc[0][i] + facc[1][i] + facc[2][i] +
                       facc[3][i];
      y[H.perm[p]] = v;
    }
    __syncthreads();
  }
}

template <typename T>
void launch_leaf(const smx_hmatrix* H, double* y, hipStream_t st) {
  if (H->near_dtype == SMX_F32) {
    hipLaunchKernelGGL((hm_leaf_kernel<T, float>), dim3((unsigned)H->nleaf), dim3(256), 0, st, *H, y);
  } else {
    hipLaunchKernelGGL((hm_leaf_kernel<T, double>), dim3((unsigned)H->nleaf), dim3(256), 0, st, *H, y);
  }
}

}  // namespace

extern "C" int smx_aca_batched(double* A, const smx_aca_task* tasks, int64_t ntasks, double rtol,
                               double* U, double* V, int32_t* rank_out, double* resid_out,
                               int32_t* status_out, void* stream) {
  SMX_REQUIRE(ntasks >= 0, SMX_ERR_SHAPE, "negative task count");
  if (ntasks == 0) return SMX_OK;
  SMX_REQUIRE(ntasks <= 0x7fffffffLL, SMX_ERR_SHAPE, "too many tasks");
  SMX_REQUIRE(A && tasks && U && V && rank_out && resid_out && status_out, SMX_ERR_VALUE,
              "null pointer argument");
  hipLaunchKernelGGL(aca_kernel, dim3((unsigned)ntasks), dim3(kAcaThreads), 0,
                     smx::as_stream(stream), A, tasks, rtol, U, V, rank_out, resid_out, status_out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_hmat_matvec(const smx_hmatrix* H, const double* src, const double* shift,
                               double shift_scale, double* y, void* stream) {
  SMX_REQUIRE(H, SMX_ERR_VALUE, "null H");
  SMX_REQUIRE(H->nb >= 0 && H->nfar >= 0 && H->nleaf >= 0, SMX_ERR_SHAPE, "negative size");
  if (H->nb == 0) return SMX_OK;
  SMX_REQUIRE(src && y && H->gidx && H->perm && H->diag_p && H->xp && H->leaf_start &&
                  H->leaf_stop && H->leaf_fe_ptr,
              SMX_ERR_VALUE, "null pointer argument");
  SMX_REQUIRE(H->factor_dtype == SMX_F32 || H->factor_dtype == SMX_F64, SMX_ERR_TYPE,
              "factor dtype must be FP32 or FP64");
  SMX_REQUIRE(H->near_dtype == SMX_F32 || H->near_dtype == SMX_F64, SMX_ERR_TYPE,
              "near dtype must be FP32 or FP64");
  SMX_REQUIRE(H->nfar == 0 || (H->far_c0 && H->far_n && H->far_m && H->far_rank && H->far_u_off &&
                               H->far_v_off && H->far_z_off && H->U && H->V && H->z),
              SMX_ERR_VALUE, "far-field arrays missing");
  hipStream_t st = smx::as_stream(stream);
  hipLaunchKernelGGL(hm_gather_kernel, dim3(smx::grid_for(H->nb, 256)), dim3(256), 0, st, H->nb,
                     src, H->gidx, shift, shift_scale, H->xp);
  SMX_LAUNCH_CHECK();
  if (H->nfar > 0) {
    const int g = smx::grid_for(H->nfar, 4);
    if (H->factor_dtype == SMX_F32)
      hipLaunchKernelGGL(hm_far_z_kernel<float>, dim3(g), dim3(256), 0, st, *H);
    else
      hipLaunchKernelGGL(hm_far_z_kernel<double>, dim3(g), dim3(256), 0, st, *H);
    SMX_LAUNCH_CHECK();
  }
  if (H->nleaf > 0) {
    if (H->factor_dtype == SMX_F32)
      launch_leaf<float>(H, y, st);
    else
      launch_leaf<double>(H, y, st);
    SMX_LAUNCH_CHECK();
  }
  return SMX_OK;
}
