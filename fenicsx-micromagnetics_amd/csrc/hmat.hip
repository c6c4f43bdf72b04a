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
#include <mutex>

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
      if (bi < 0 || bi >= (int64_t)m * n) {  // no valid pivot (a block of NaNs)
        status = 1;
        break;
      }
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


// Large blocks: the same full-pivot ACA, each rank step spread over row
// tiles of kAcaTileRows rows (one workgroup per tile, the tiles of every block
// of the call in one launch), then one pivot workgroup per block.  Per
// element the update is the single-workgroup kernel's (unfused multiply and
// subtract); the first-occurrence argmax does not depend on the tiling, and
// ||R||_F^2 is summed per tile, then over the tiles in order.
// Per-block state (8 doubles): normA, rel, done, rank, status, pi, pj, piv.
constexpr int kAcaTileRows = SMX_ACA_TILE_ROWS;
enum { AS_NORM = 0, AS_REL, AS_DONE, AS_RANK, AS_STATUS, AS_PI, AS_PJ, AS_PIV, AS_SIZE };

// Step k of every open block: apply R -= u_{k-1} v_{k-1}^T to the tile's rows
// (k > 0), then the tile's (sum of squares, argmax |R|) partials.
__global__ __launch_bounds__(kAcaThreads) void aca_tile_kernel(double* __restrict__ A,
                                                               const smx_aca_task* __restrict__ tasks,
                                                               const int32_t* __restrict__ tile_task,
                                                               const int32_t* __restrict__ tile_row0,
                                                               const double* __restrict__ U,
                                                               const double* __restrict__ V,
                                                               const double* __restrict__ state,
                                                               double* __restrict__ part, int32_t k) {
  __shared__ double ss[kAcaWaves + 1];
  __shared__ double sv[kAcaWaves + 1];
  __shared__ int64_t si[kAcaWaves + 1];
  const int b = tile_task[blockIdx.x];
  if (state[(int64_t)b * AS_SIZE + AS_DONE] != 0.0) return;  // uniform
  const smx_aca_task T = tasks[b];
  const int m = T.m, n = T.n;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int r0 = tile_row0[blockIdx.x];
  const int r1 = r0 + kAcaTileRows < m ? r0 + kAcaTileRows : m;
  double* __restrict__ R = A + T.a_off;
  const double* __restrict__ u = k > 0 ? U + T.u_off + (int64_t)(k - 1) * m : nullptr;
  const double* __restrict__ v = k > 0 ? V + T.v_off + (int64_t)(k - 1) * n : nullptr;
  double s = 0.0, bv = -1.0;
  int64_t bi = INT64_MAX;
  for (int r = r0 + wid; r < r1; r += kAcaWaves) {
    double* row = R + (int64_t)r * n;
    const double ur = u ? u[r] : 0.0;
    for (int c = lane; c < n; c += 64) {
      double nv = row[c];
      if (u) {
        nv = __dsub_rn(nv, __dmul_rn(ur, v[c]));
        row[c] = nv;
      }
      s += nv * nv;
      am_merge(bv, bi, fabs(nv), (int64_t)r * n + c);
    }
  }
  block_sum_argmax(s, bv, bi, ss, sv, si);
  if (threadIdx.x == 0) {
    part[3 * (int64_t)blockIdx.x] = s;
    part[3 * (int64_t)blockIdx.x + 1] = bv;
    part[3 * (int64_t)blockIdx.x + 2] = (double)bi;
  }
}

// Step k of block b: the tiles' partials reduced in tile order, the stopping
// tests of aca_kernel in its order (zero block; rel <= rtol after k updates;
// k == rcap; pivot below eps ||A||_F), and, for an open block, u_k = R[:, pj]
// and v_k = R[pi, :] / pivot into U and V.  A block that stops writes its
// rank, residual and status and counts itself in done_count.
__global__ __launch_bounds__(kAcaThreads) void aca_pivot_kernel(double* __restrict__ A,
                                                                const smx_aca_task* __restrict__ tasks,
                                                                const int32_t* __restrict__ tile_first,
                                                                double rtol, double* __restrict__ U,
                                                                double* __restrict__ V,
                                                                double* __restrict__ state,
                                                                const double* __restrict__ part,
                                                                int32_t k, int32_t* rank_out,
                                                                double* resid_out, int32_t* status_out,
                                                                int32_t* done_count) {
  __shared__ double ss[kAcaWaves + 1];
  __shared__ double sv[kAcaWaves + 1];
  __shared__ int64_t si[kAcaWaves + 1];
  __shared__ double dec[4];
  const int b = blockIdx.x;
  double* st = state + (int64_t)b * AS_SIZE;
  if (st[AS_DONE] != 0.0) return;  // uniform
  const smx_aca_task T = tasks[b];
  const int m = T.m, n = T.n;
  double* __restrict__ R = A + T.a_off;
  double s = 0.0, bv = -1.0;
  int64_t bi = INT64_MAX;
  const int t0 = tile_first[b], t1 = tile_first[b + 1];
  for (int t = t0 + (int)threadIdx.x; t < t1; t += kAcaThreads) {
    s += part[3 * (int64_t)t];
    am_merge(bv, bi, part[3 * (int64_t)t + 1], (int64_t)part[3 * (int64_t)t + 2]);
  }
  block_sum_argmax(s, bv, bi, ss, sv, si);
  if (threadIdx.x == 0) {
    double normA = st[AS_NORM], rel = k == 0 ? 1.0 : st[AS_REL];
    int rank = k, status = -1;
    if (k == 0) {
      normA = sqrt(s);
      st[AS_NORM] = normA;
      if (normA == 0.0 || T.rcap == 0) {
        status = 3;
        rel = 0.0;
      }
    } else {
      rel = sqrt(s) / normA;
      if (rel <= rtol) status = 0;
    }
    if (status < 0 && k == T.rcap) status = 2;
    double piv = 0.0;
    if (status < 0) {
      // no valid pivot (a block of NaNs) stops as a pivot break
      piv = (bi >= 0 && bi < (int64_t)m * n) ? R[bi] : 0.0;
      if (!(fabs(piv) > DBL_EPSILON * normA)) status = 1;
    }
    st[AS_REL] = rel;
    st[AS_RANK] = (double)rank;
    if (status >= 0) {
      st[AS_STATUS] = (double)status;
      st[AS_DONE] = 1.0;
      rank_out[b] = rank;
      resid_out[b] = rel;
      status_out[b] = status;
      atomicAdd(done_count, 1);
    }
    dec[0] = status >= 0 ? 1.0 : 0.0;
    dec[1] = (double)(bi / n);
    dec[2] = (double)(bi - (bi / n) * n);
    dec[3] = piv;
  }
  __syncthreads();
  if (dec[0] != 0.0) return;  // uniform
  const int pi = (int)dec[1], pj = (int)dec[2];
  const double piv = dec[3];
  double* __restrict__ u = U + T.u_off + (int64_t)k * m;
  double* __restrict__ v = V + T.v_off + (int64_t)k * n;
  for (int r = threadIdx.x; r < m; r += kAcaThreads) u[r] = R[(int64_t)r * n + pj];
  for (int c = threadIdx.x; c < n; c += kAcaThreads) v[c] = R[(int64_t)pi * n + c] / piv;
}

__global__ void aca_state_init_kernel(double* __restrict__ state, int64_t nblocks, int32_t* done_count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nblocks * AS_SIZE) state[i] = 0.0;
  if (i == 0) *done_count = 0;
}

__global__ void hm_gather_kernel(int64_t nb, const double* __restrict__ src,
                                 const int32_t* __restrict__ gidx, const double* __restrict__ shift,
                                 double scale, double* __restrict__ xp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const double sh = shift ? (*shift) * scale : 0.0;
  xp[i] = src[gidx[i]] - sh;
}

// Streamed operands (FP32 factors, the near field) are loaded with the
// non-temporal hint, so they do not push the reused x and z slices out of L2:
// film matvec 326 -> 308 us (FP32 factors).  The FP64 factors keep plain
// loads: with the hint their z launch slows from 245 to 310 us.
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef int nt_i2 __attribute__((ext_vector_type(2)));

// Four consecutive factor entries (16-byte aligned) widened to double.
__device__ __forceinline__ void load4(const float* __restrict__ p, double (&o)[4]) {
  const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
  o[0] = v.x;
  o[1] = v.y;
  o[2] = v.z;
  o[3] = v.w;
}
__device__ __forceinline__ void load4(const double* __restrict__ p, double (&o)[4]) {
  const double2 a = *reinterpret_cast<const double2*>(p);
  const double2 b = *reinterpret_cast<const double2*>(p + 2);
  o[0] = a.x;
  o[1] = a.y;
  o[2] = b.x;
  o[3] = b.y;
}
// A pair of near-field entries and their columns (non-temporal).
__device__ __forceinline__ void load2(const float* __restrict__ p, double& a, double& b) {
  const nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2*>(p));
  a = v.x;
  b = v.y;
}
__device__ __forceinline__ void load2(const double* __restrict__ p, double& a, double& b) {
  const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
  a = v.x;
  b = v.y;
}
__device__ __forceinline__ int2 load_cols2(const int32_t* __restrict__ p) {
  const nt_i2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_i2*>(p));
  return make_int2(v.x, v.y);
}

constexpr int kZG = SMX_HMAT_Z_GROUP;
static_assert(kZG == 8, "the reduce-scatter below is written for 8 ranks per task");

// z task: kc rank rows of one far block against its column slice of xp.
// Lanes stride the columns four at a time (16-byte factor loads); each lane
// keeps kZG partial sums, then a butterfly reduce-scatter (xor 32, 16, 8 halve
// the set, xor 4, 2, 1 finish the sums) leaves rank g's total in lanes whose
// bits 5..3 spell g.  Fixed order, so the result is deterministic.  The task
// record holds everything the wavefront needs in one load.
//
// The two factor types take different forms, each the faster one measured on
// the film's 74 379 tasks (one in five has kc < 8):
//  - FP32: the task index goes through readfirstlane, so the record is a
//    scalar load and rows past kc are skipped by a scalar branch: 127 us
//    against 146 us for the branch-free body that re-reads row kc-1;
//  - FP64: eight rows of four doubles fill the 64-VGPR budget of 8 waves per
//    SIMD.  With a scalar record the compiler serializes the row loads
//    (290-470 us), so the record stays per-lane and the body stays branch-free
//    with row kc-1 re-read (250 us).
// Near rows in a launch of their own: one wavefront per row,
// ynear[p] = Near[p, :] . xp (pairs of entries per lane, as the leaf kernel's
// near loop), written by lane 0.  Spreads the near field over the whole
// device when the leaves are too few to fill it.
template <typename TN>
__global__ __launch_bounds__(256) void hm_near_kernel(smx_hmatrix H) {
  const int64_t p = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= H.nb) return;  // wave-uniform; no barrier
  const TN* __restrict__ NV = reinterpret_cast<const TN*>(H.near_val);
  const double* __restrict__ xp = H.xp;
  double s = 0.0;
  const int64_t n1 = H.near_ptr[p + 1];
  for (int64_t e = H.near_ptr[p] + 2 * lane; e < n1; e += 128) {
    double v0, v1;
    load2(NV + e, v0, v1);
    const int2 c = load_cols2(H.near_idx + e);
    s = fma(v0, xp[c.x], s);
    s = fma(v1, xp[c.y], s);
  }
  s = smx::wave_total(s);
  if (lane == 0) H.ynear[p] = s;
}

// Near rows by task (one workgroup per task): the task's columns of xp are
// staged in LDS once, then each wavefront sums whole rows as hm_near_kernel
// does (the same entries per lane, the same fma order, the same wave total),
// reading a 2-byte slot per entry instead of a 4-byte column and x from LDS
// instead of the L2.
template <typename TN, int NW>
__global__ __launch_bounds__(NW * 64) void hm_near_lds_kernel(smx_hmatrix H) {
  extern __shared__ double xs[];
  const smx_hmat_ntask T = H.near_task[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int32_t* __restrict__ stg = H.near_stage + T.s_off;
  for (int i = threadIdx.x; i < T.w; i += NW * 64) xs[i] = H.xp[stg[i]];
  __syncthreads();
  const TN* __restrict__ NV = reinterpret_cast<const TN*>(H.near_val);
  const uint32_t* __restrict__ LI = reinterpret_cast<const uint32_t*>(H.near_lidx);
  for (int64_t p = T.p0 + wid; p < T.p1; p += NW) {
    double s = 0.0;
    const int64_t n1 = H.near_ptr[p + 1];
    for (int64_t e = H.near_ptr[p] + 2 * lane; e < n1; e += 128) {
      double v0, v1;
      load2(NV + e, v0, v1);
      const uint32_t c = __builtin_nontemporal_load(LI + (e >> 1));  // slots e, e+1 (e even)
      s = fma(v0, xs[c & 0xffffu], s);
      s = fma(v1, xs[c >> 16], s);
    }
    s = smx::wave_total(s);
    if (lane == 0) H.ynear[p] = s;
  }
}

template <typename T>
__global__ __launch_bounds__(256, 8) void hm_far_z_kernel(smx_hmatrix H) {
  constexpr bool kF32 = sizeof(T) == 4;
  const int wv = (int)threadIdx.x >> 6;
  const int64_t t = (int64_t)blockIdx.x * 4 + (kF32 ? __builtin_amdgcn_readfirstlane(wv) : wv);
  const int lane = threadIdx.x & 63;
  if (t >= H.nztask) return;  // wave-uniform; the kernel has no barrier
  const smx_hmat_ztask task = H.ztask[t];
  const int32_t n = task.n;
  const int kc = task.kc;
  constexpr int64_t kA = sizeof(T) == 4 ? SMX_HMAT_V_ALIGN_F32 : SMX_HMAT_V_ALIGN_F64;  // V row padding
  const int64_t ld = ((int64_t)n + kA - 1) & ~(kA - 1);
  const T* __restrict__ Vt = reinterpret_cast<const T*>(H.V) + task.v_off;
  const double* __restrict__ x = H.xp + task.c0;
  double acc[kZG];
#pragma unroll
  for (int g = 0; g < kZG; ++g) acc[g] = 0.0;
  for (int j = 4 * lane; j < n; j += 256) {
    double xv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) xv[q] = x[(j + q) < n ? (j + q) : (n - 1)];  // padded V entries are 0
#pragma unroll
    for (int g = 0; g < kZG; ++g) {
      int gg = g;
      if constexpr (kF32) {
        if (g >= kc) break;
      } else {
        gg = g < kc ? g : kc - 1;  // re-read row kc-1; the sums of rows past kc are dropped
      }
      double v[4];
      load4(Vt + (int64_t)gg * ld + j, v);
      acc[g] = fma(v[0], xv[0], acc[g]);
      acc[g] = fma(v[1], xv[1], acc[g]);
      acc[g] = fma(v[2], xv[2], acc[g]);
      acc[g] = fma(v[3], xv[3], acc[g]);
    }
  }
  const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
  double a4[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    a4[i] = (h5 ? acc[4 + i] : acc[i]) + __shfl_xor(h5 ? acc[i] : acc[4 + i], 32, 64);
  double a2[2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
    a2[i] = (h4 ? a4[2 + i] : a4[i]) + __shfl_xor(h4 ? a4[i] : a4[2 + i], 16, 64);
  double a1 = (h3 ? a2[1] : a2[0]) + __shfl_xor(h3 ? a2[0] : a2[1], 8, 64);
  a1 += __shfl_xor(a1, 4, 64);
  a1 += __shfl_xor(a1, 2, 64);
  a1 += __shfl_xor(a1, 1, 64);
  const int g = (h5 ? 4 : 0) + (h4 ? 2 : 0) + (h3 ? 1 : 0);
  if ((lane & 7) == 0 && g < kc) H.z[task.z_off + g] = a1;
}

// One workgroup per leaf (leaves are stored largest first, so the dispatcher
// balances the tail): the leaf's z entries are gathered into LDS in column
// order, then each wavefront produces whole output rows as one dot product
// (near pairs against xp, then the packed far row against the staged z).
template <typename T, typename TN, int NW>
__global__ __launch_bounds__(NW * 64) void hm_leaf_kernel(smx_hmatrix H, double* __restrict__ y) {
  extern __shared__ double zl[];
  const int64_t L = blockIdx.x;
  const int lane = threadIdx.x & 63;
  // wave-uniform row loop (readfirstlane), so the row's wave total can use
  // DPP lane moves and readlane: 2 % faster than shuffles on the film (FP32)
  const int wid = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int32_t l0 = H.leaf_start[L], l1 = H.leaf_stop[L];
  const int32_t kw = H.leaf_kw[L];
  const int64_t g0 = H.leaf_zg_off[L];
  const int32_t K = (int32_t)(H.leaf_zg_off[L + 1] - g0);
  if (H.leaf_zr) {  // contiguous runs of z (uniform loop, scalar run records)
    const int64_t q1 = H.leaf_zr_off[L + 1];
    for (int64_t q = H.leaf_zr_off[L]; q < q1; ++q) {
      const int32_t zs = H.leaf_zr[2 * q], k0 = H.leaf_zr[2 * q + 1];
      const int32_t k1 = q + 1 < q1 ? H.leaf_zr[2 * q + 3] : K;
      for (int k = k0 + (int)threadIdx.x; k < k1; k += NW * 64) zl[k] = H.z[zs + (k - k0)];
    }
    for (int k = K + (int)threadIdx.x; k < kw; k += NW * 64) zl[k] = 0.0;
  } else {
    const int32_t* __restrict__ zg = H.leaf_zg + g0;
    for (int k = threadIdx.x; k < kw; k += NW * 64) zl[k] = k < K ? H.z[zg[k]] : 0.0;
  }
  __syncthreads();
  const T* __restrict__ UL = reinterpret_cast<const T*>(H.U) + H.leaf_u_off[L];
  const TN* __restrict__ NV = reinterpret_cast<const TN*>(H.near_val);
  const double* __restrict__ xp = H.xp;
  for (int32_t i = wid; i < l1 - l0; i += NW) {
    const int64_t p = l0 + i;
    double s = 0.0;
    if (H.near_ptr && !H.ynear) {
      const int64_t n1 = H.near_ptr[p + 1];
      for (int64_t e = H.near_ptr[p] + 2 * lane; e < n1; e += 128) {
        double v0, v1;
        load2(NV + e, v0, v1);
        const int2 c = load_cols2(H.near_idx + e);
        s = fma(v0, xp[c.x], s);
        s = fma(v1, xp[c.y], s);
      }
    }
    const T* __restrict__ Ur = UL + (int64_t)i * kw;
    for (int k = 4 * lane; k < kw; k += 256) {
      double u[4];
      load4(Ur + k, u);
      s = fma(u[0], zl[k], s);
      s = fma(u[1], zl[k + 1], s);
      s = fma(u[2], zl[k + 2], s);
      s = fma(u[3], zl[k + 3], s);
    }
    s = smx::wave_total(s);
    if (H.ynear && H.near_ptr) s += H.ynear[p];  // near row summed in the z launch
    if (lane == 0) y[H.perm[p]] = H.diag_p[p] * xp[p] + s;
  }
}

constexpr int kLeafWaves = 8;

template <typename T, int NW>
void launch_leaf_nw(const smx_hmatrix* H, double* y, hipStream_t st, size_t lds) {
  if (H->near_dtype == SMX_F32) {
    hipLaunchKernelGGL((hm_leaf_kernel<T, float, NW>), dim3((unsigned)H->nleaf), dim3(NW * 64), lds, st, *H, y);
  } else {
    hipLaunchKernelGGL((hm_leaf_kernel<T, double, NW>), dim3((unsigned)H->nleaf), dim3(NW * 64), lds, st, *H,
                       y);
  }
}

// leaf_waves: wavefronts per leaf workgroup (4, 8 or 16); 0 picks the
// measured best on the film's 2048 leaves: 8 for FP32 factors (326 us
// against 333 with 16 and 352 with 4), 16 for FP64 (543 against 550 with 8)
template <typename T>
void launch_leaf(const smx_hmatrix* H, double* y, hipStream_t st) {
  const size_t lds = sizeof(double) * (size_t)(H->kw_max > 0 ? H->kw_max : 1);
  const int nw = H->leaf_waves ? H->leaf_waves : (sizeof(T) == 8 ? 16 : kLeafWaves);
  switch (nw) {
    case 4: launch_leaf_nw<T, 4>(H, y, st, lds); break;
    case 16: launch_leaf_nw<T, 16>(H, y, st, lds); break;
    default: launch_leaf_nw<T, kLeafWaves>(H, y, st, lds); break;
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


extern "C" int64_t smx_aca_large_ws_doubles(int64_t ntasks, int64_t ntiles) {
  if (ntasks < 0 || ntiles < 0) return -1;
  return ntasks * AS_SIZE + 3 * ntiles + 2;
}

extern "C" int smx_aca_large(double* A, const smx_aca_task* tasks, int64_t ntasks, const int32_t* tile_task,
                             const int32_t* tile_row0, const int32_t* tile_first, int64_t ntiles,
                             int32_t max_rcap, double rtol, double* U, double* V, int32_t* rank_out,
                             double* resid_out, int32_t* status_out, double* work, int64_t work_doubles,
                             void* stream) {
  SMX_REQUIRE(ntasks >= 0 && ntiles >= 0 && max_rcap >= 0, SMX_ERR_SHAPE, "negative size");
  if (ntasks == 0) return SMX_OK;
  SMX_REQUIRE(ntasks <= 0x7fffffffLL && ntiles <= 0x7fffffffLL, SMX_ERR_SHAPE, "too many tasks or tiles");
  SMX_REQUIRE(A && tasks && tile_task && tile_row0 && tile_first && U && V && rank_out && resid_out &&
                  status_out && work,
              SMX_ERR_VALUE, "null pointer argument");
  SMX_REQUIRE(work_doubles >= smx_aca_large_ws_doubles(ntasks, ntiles), SMX_ERR_OOM,
              "workspace of %lld doubles, %lld needed", (long long)work_doubles,
              (long long)smx_aca_large_ws_doubles(ntasks, ntiles));
  hipStream_t st = smx::as_stream(stream);
  double* state = work;
  double* part = work + ntasks * AS_SIZE;
  int32_t* done = reinterpret_cast<int32_t*>(part + 3 * ntiles);
  const int64_t ninit = ntasks * AS_SIZE;
  hipLaunchKernelGGL(aca_state_init_kernel, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, st, state,
                     ntasks, done);
  SMX_LAUNCH_CHECK();
  int32_t hdone = 0;
  constexpr int kCheckEvery = 8;
  for (int32_t k = 0; k <= max_rcap; ++k) {
    if (ntiles > 0) {
      hipLaunchKernelGGL(aca_tile_kernel, dim3((unsigned)ntiles), dim3(kAcaThreads), 0, st, A, tasks, tile_task,
                         tile_row0, U, V, state, part, k);
      SMX_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(aca_pivot_kernel, dim3((unsigned)ntasks), dim3(kAcaThreads), 0, st, A, tasks, tile_first,
                       rtol, U, V, state, part, k, rank_out, resid_out, status_out, done);
    SMX_LAUNCH_CHECK();
    if ((k + 1) % kCheckEvery == 0 || k == max_rcap) {
      SMX_HIP_CHECK(hipMemcpyAsync(&hdone, done, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      SMX_HIP_CHECK(hipStreamSynchronize(st));
      if (hdone >= ntasks) break;
    }
  }
  SMX_REQUIRE(hdone >= ntasks, SMX_ERR_VALUE, "%d of %lld blocks still open after %d steps (max_rcap too small)",
              hdone, (long long)ntasks, max_rcap + 1);
  return SMX_OK;
}

namespace {

int hmat_matvec_impl(const smx_hmatrix* H, const double* src, const double* shift,
                     double shift_scale, double* y, void* stream) {
  SMX_REQUIRE(H, SMX_ERR_VALUE, "null H");
  SMX_REQUIRE(H->nb >= 0 && H->nleaf >= 0 && H->nztask >= 0, SMX_ERR_SHAPE, "negative size");
  if (H->nb == 0) return SMX_OK;
  SMX_REQUIRE(src && y && H->gidx && H->perm && H->diag_p && H->xp && H->leaf_start &&
                  H->leaf_stop && H->leaf_kw && H->leaf_u_off && H->leaf_zg_off && H->U,
              SMX_ERR_VALUE, "null pointer argument");
  SMX_REQUIRE(H->factor_dtype == SMX_F32 || H->factor_dtype == SMX_F64, SMX_ERR_TYPE,
              "factor dtype must be FP32 or FP64");
  SMX_REQUIRE(H->near_dtype == SMX_F32 || H->near_dtype == SMX_F64, SMX_ERR_TYPE,
              "near dtype must be FP32 or FP64");
  SMX_REQUIRE(!H->leaf_zr == !H->leaf_zr_off, SMX_ERR_VALUE, "leaf_zr and leaf_zr_off go together");
  SMX_REQUIRE(H->nztask == 0 || (H->ztask && H->V && H->z && H->leaf_zg), SMX_ERR_VALUE,
              "far-field arrays missing");
  SMX_REQUIRE(H->leaf_waves == 0 || H->leaf_waves == 4 || H->leaf_waves == 8 || H->leaf_waves == 16,
              SMX_ERR_VALUE, "leaf_waves must be 0, 4, 8 or 16");
  SMX_REQUIRE(H->kw_max >= 0 && H->kw_max <= SMX_HMAT_KW_MAX && (H->kw_max & 3) == 0, SMX_ERR_SHAPE,
              "kw_max %d must be a multiple of 4 in [0, %d]", H->kw_max, SMX_HMAT_KW_MAX);
  SMX_REQUIRE(H->nleaf <= 0x7fffffffLL && H->nztask / 4 < 0x7fffffffLL && H->nb / 4 < 0x7fffffffLL, SMX_ERR_SHAPE,
              "too many leaves or z tasks");
  SMX_REQUIRE(!H->ynear || !H->near_ptr || (H->near_idx && H->near_val), SMX_ERR_VALUE,
              "near arrays missing");
  const bool near_tasks = H->ynear && H->near_ptr && H->near_task;
  SMX_REQUIRE(!near_tasks || (H->near_stage && H->near_lidx && H->n_near_task > 0 &&
                              H->n_near_task <= 0x7fffffffLL && H->near_w_max >= 0 &&
                              H->near_w_max <= SMX_HMAT_NEAR_W_MAX),
              SMX_ERR_VALUE, "near task arrays missing or near_w_max beyond %d", SMX_HMAT_NEAR_W_MAX);
  SMX_REQUIRE(H->near_waves == 0 || H->near_waves == 4 || H->near_waves == 8 || H->near_waves == 16,
              SMX_ERR_VALUE, "near_waves must be 0, 4, 8 or 16");
  hipStream_t st = smx::as_stream(stream);
  hipLaunchKernelGGL(hm_gather_kernel, dim3(smx::grid_for(H->nb, 256)), dim3(256), 0, st, H->nb,
                     src, H->gidx, shift, shift_scale, H->xp);
  SMX_LAUNCH_CHECK();
  if (near_tasks) {
    const size_t lds = (size_t)H->near_w_max * sizeof(double);
    const unsigned gn = (unsigned)H->n_near_task;
    const int nw = H->near_waves == 0 ? 8 : H->near_waves;  // 8: fastest on the 83k-node film
#define SMX_NEAR_LDS(TN, NW) \
  hipLaunchKernelGGL((hm_near_lds_kernel<TN, NW>), dim3(gn), dim3(NW * 64), lds, st, *H)
    if (H->near_dtype == SMX_F32) {
      if (nw == 4) SMX_NEAR_LDS(float, 4); else if (nw == 16) SMX_NEAR_LDS(float, 16); else SMX_NEAR_LDS(float, 8);
    } else {
      if (nw == 4) SMX_NEAR_LDS(double, 4); else if (nw == 16) SMX_NEAR_LDS(double, 16); else SMX_NEAR_LDS(double, 8);
    }
#undef SMX_NEAR_LDS
    SMX_LAUNCH_CHECK();
  } else if (H->ynear && H->near_ptr) {
    const unsigned gn = (unsigned)((H->nb + 3) / 4);
    if (H->near_dtype == SMX_F32)
      hipLaunchKernelGGL(hm_near_kernel<float>, dim3(gn), dim3(256), 0, st, *H);
    else
      hipLaunchKernelGGL(hm_near_kernel<double>, dim3(gn), dim3(256), 0, st, *H);
    SMX_LAUNCH_CHECK();
  }
  if (H->nztask > 0) {
    const unsigned g = (unsigned)((H->nztask + 3) / 4);
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

// Scratch of the entry point that allocates nothing: xp, z and ynear of
// smx_hmat_matvec live in this array instead of caller buffers.  Each GPU the
// library runs on holds one copy until the process ends.
constexpr int64_t kScratchDoubles = SMX_HMAT_NOALLOC_MAX_DOUBLES;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (the host-side name of a __device__ array is not its address).
int scratch_ptr(double** out) {
  int dev = 0;
  SMX_HIP_CHECK(hipGetDevice(&dev));
  SMX_REQUIRE(dev >= 0 && dev < kMaxDevices, SMX_ERR_VALUE, "device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SMX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SMX_REQUIRE(p, SMX_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SMX_OK;
}

}  // namespace

extern "C" int smx_hmat_matvec(const smx_hmatrix* H, const double* src, const double* shift,
                               double shift_scale, double* y, void* stream) {
  return hmat_matvec_impl(H, src, shift, shift_scale, y, stream);
}

extern "C" int smx_hmat_matvec_noalloc(const smx_hmatrix* H, int64_t nz, const double* src,
                                       const double* shift, double shift_scale, double* y,
                                       void* stream) {
  SMX_REQUIRE(H, SMX_ERR_VALUE, "null H");
  SMX_REQUIRE(H->nb >= 0 && nz >= 0, SMX_ERR_SHAPE, "negative size");
  const int64_t need = H->nb + nz + (H->ynear ? H->nb : 0);
  SMX_REQUIRE(need <= kScratchDoubles, SMX_ERR_SHAPE,
              "smx_hmat_matvec_noalloc: the operator needs %lld scratch doubles (nb %lld, %lld z "
              "entries%s), more than the %lld of the fixed scratch array; use smx_hmat_matvec with "
              "caller workspaces", (long long)need, (long long)H->nb, (long long)nz,
              H->ynear ? ", near rows in a launch of their own" : "", (long long)kScratchDoubles);
  if (H->nb == 0) return SMX_OK;
  double* p = nullptr;
  const int rc = scratch_ptr(&p);
  if (rc) return rc;
  smx_hmatrix Hs = *H;
  Hs.xp = p;
  Hs.z = p + H->nb;
  Hs.ynear = H->ynear ? p + H->nb + nz : nullptr;
  return hmat_matvec_impl(&Hs, src, shift, shift_scale, y, stream);
}
