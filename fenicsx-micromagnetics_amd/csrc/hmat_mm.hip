// Multi-RHS H-matrix product Y = B X (up to 64 right-hand sides) with the
// low-rank products on the matrix cores.
//
// Same packed operator as smx_hmat_matvec (z tasks over V, leaves over U,
// near CSR); with 64 columns of X the far-field work becomes small GEMMs:
//   stage:  XR[p][j] = X[j][gidx[p]] in the factor type, p < nb, j < 64
//           (and XD, the same in FP64)
//   z:      Z[z_off + g][j] = V_t[g, :] . XR[c0 .. c0+n)[j]      (MFMA)
//   leaf:   YF[p][j] = U_L[p - l0, :] . Z[zg_L[k]][j]           (MFMA)
//   near:   Y[j][perm[p]] = d_p XR[p][j] + Near[p,:] XR[:, j] + YF[p][j]
//           (dense 16-row near tiles on the FP64 matrix cores, or CSR rows)
// Every operand array is row-major with a row's 64 columns contiguous; XR
// (and Z for FP64 factors) interleave them (column 16 q + c at 4 c + q), so
// the 16 lanes of a lane group read one whole row, each lane its four column
// blocks at once.
// MFMA per factor type (lane l, g = l / 16, c = l % 16), as the dense
// multi-RHS product: FP32 v_mfma_f32_16x16x4_f32, D[m = 4g + i][n = c];
// FP64 v_mfma_f64_16x16x4_f64, D[m = g + 4i][n = c]; operand A[m = c][k],
// B[k][n = c], the k of MFMA j in a step of lane group g being
// kk + KL g + j for both operands.  A z task fills rows 0..kc-1 of a
// 16-row tile (rows past kc load zeros); with FP32 factors a leaf
// workgroup's four wavefronts split its columns and each covers all of its
// 16-row tiles (hmm_leaf_ks_kernel), with FP64 they cover one tile each.  FP32 factors: X, Z in FP32 (one rounding
// each), FP32 accumulation in the tiles; the near field (over an FP64 copy of
// X) and all sums in FP64.  Z tasks and leaves run largest first; the
// near-tile launch maps workgroups to XCDs by contiguous eighths of its
// (surface-ordered) tiles.
#include "smx_common.h"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kR = 64;  // right-hand sides per launch (columns of XR, Z, YF)
constexpr int kNtCols = 1024;  // near-tile columns staged in LDS at a time (a multiple of 64)
constexpr int kLeafCols = 2048;  // leaf Z indices staged in LDS at a time (a multiple of 32)

template <typename T>
struct Mt;

// FP32: a step is 32 k; a lane loads 8 consecutive k = 8 MFMAs of 4.
template <>
struct Mt<float> {
  using acc_t = f32x4;
  static constexpr int kStep = 32;
  static constexpr int kLane = 8;
  // Z in column order: the leaf stage's eight Z rows per step gather best
  // as single 4-byte loads per column block (measured on the film)
  static constexpr bool kZInterleaved = false;
  // 8 factor entries from p, `room` entries (a multiple of 4) left in the row
  __device__ static void load_a(const float* p, int64_t room, float (&v)[8]) {
    if (room >= 4) {
      const float4 a = *reinterpret_cast<const float4*>(p);
      v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
    } else {
      v[0] = v[1] = v[2] = v[3] = 0.0f;
    }
    if (room >= 8) {
      const float4 b = *reinterpret_cast<const float4*>(p + 4);
      v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
    } else {
      v[4] = v[5] = v[6] = v[7] = 0.0f;
    }
  }
  __device__ static void mma(acc_t& c, const float (&a)[8], const float (&b)[8]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j], b[j], c, 0, 0, 0);
  }
  __device__ static int row_of(int g, int i) { return 4 * g + i; }
};

// FP64: a step is 16 k; a lane loads 4 consecutive k = 4 MFMAs of 4.
template <>
struct Mt<double> {
  using acc_t = f64x4;
  static constexpr int kStep = 16;
  static constexpr int kLane = 4;
  static constexpr bool kZInterleaved = true;
  __device__ static void load_a(const double* p, int64_t room, double (&v)[4]) {
    if (room >= 4) {
      const double2 a = *reinterpret_cast<const double2*>(p);
      const double2 b = *reinterpret_cast<const double2*>(p + 2);
      v[0] = a.x, v[1] = a.y, v[2] = b.x, v[3] = b.y;
    } else {
      v[0] = v[1] = v[2] = v[3] = 0.0;
    }
  }
  __device__ static void mma(acc_t& c, const double (&a)[4], const double (&b)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j], b[j], c, 0, 0, 0);
  }
  __device__ static int row_of(int g, int i) { return g + 4 * i; }
};

// XR (and Z for FP64 factors) keep a row's 64 columns interleaved: column
// 16 q + c at 4 c + q,
// so lane c of a lane group reads (or writes) its four column blocks
// q = 0..3 as one 16-byte (FP32) or two 16-byte (FP64) accesses.
__device__ inline void load_q4(const float* p, float (&v)[4]) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
}
__device__ inline void load_q4(const double* p, double (&v)[4]) {
  const double2 a = *reinterpret_cast<const double2*>(p);
  const double2 b = *reinterpret_cast<const double2*>(p + 2);
  v[0] = a.x, v[1] = a.y, v[2] = b.x, v[3] = b.y;
}
__device__ inline void store_q4(float* p, float v0, float v1, float v2, float v3) {
  *reinterpret_cast<float4*>(p) = make_float4(v0, v1, v2, v3);
}
__device__ inline void store_q4(double* p, double v0, double v1, double v2, double v3) {
  *reinterpret_cast<double2*>(p) = make_double2(v0, v1);
  *reinterpret_cast<double2*>(p + 2) = make_double2(v2, v3);
}

// One workgroup per 64 permuted rows: the wavefronts read right-hand sides
// j = w, w + 4, ... for 64 consecutive rows (lanes along p, through gidx)
// into LDS, then write each row whole: XD in column order, XR interleaved
// (slot r holds column 16 (r mod 4) + r / 4) in the factor type.
template <typename T>
__global__ __launch_bounds__(256) void hmm_stage_kernel(int64_t nb, int32_t nrhs, const double* __restrict__ X,
                                                        int64_t ldx, const int32_t* __restrict__ gidx,
                                                        T* __restrict__ XR, double* __restrict__ XD) {
  __shared__ double tb[64][kR + 1];
  const int64_t p0 = (int64_t)blockIdx.x * 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t pl = p0 + lane;
  const int64_t src = pl < nb ? (int64_t)gidx[pl] : 0;
  for (int j = wv; j < kR; j += 4) tb[lane][j] = (j < nrhs && pl < nb) ? X[(int64_t)j * ldx + src] : 0.0;
  __syncthreads();
  const int jr = 16 * (lane & 3) + (lane >> 2);
  for (int r = wv; r < 64; r += 4) {
    const int64_t p = p0 + r;
    if (p >= nb) break;
    XD[p * kR + lane] = tb[r][lane];
    XR[p * kR + lane] = (T)tb[r][jr];
  }
}

// Workgroup b of a launch over n items runs on XCD b % 8; item
// (b % 8) * ceil(n / 8) + b / 8 gives each XCD a contiguous eighth of the
// items (stored in surface order), so its L2 serves one part of the surface.
__device__ inline int64_t xcd_item(int64_t n) {
  const int64_t per = (n + 7) / 8;
  return (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
}
inline unsigned xcd_grid(int64_t n) { return (unsigned)((n + 7) / 8 * 8); }

// One wavefront per z task; four per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void hmm_z_kernel(smx_hmatrix H, const T* __restrict__ XR,
                                                    T* __restrict__ Z) {
  using M = Mt<T>;
  const int lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  if (t >= H.nztask) return;  // wave-uniform; no barrier
  const int g = lane >> 4, c = lane & 15;
  const smx_hmat_ztask task = H.ztask[t];
  const int64_t n = task.n;
  const int kc = task.kc;
  constexpr int64_t kA = sizeof(T) == 4 ? SMX_HMAT_V_ALIGN_F32 : SMX_HMAT_V_ALIGN_F64;  // V row padding
  const int64_t ld = (n + kA - 1) & ~(kA - 1);
  const bool row_ok = c < kc;
  const T* __restrict__ Vr = reinterpret_cast<const T*>(H.V) + task.v_off + (int64_t)(row_ok ? c : 0) * ld;
  const T* __restrict__ xs = XR + (int64_t)task.c0 * kR + 4 * c;  // lane c's columns of the x segment
  typename M::acc_t acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[q][i] = 0;
  for (int64_t kk = 0; kk < n; kk += M::kStep) {
    const int64_t kl = kk + M::kLane * g;
    T a[M::kLane];
    M::load_a(Vr + kl, row_ok ? ld - kl : 0, a);
    T b[4][M::kLane];
#pragma unroll
    for (int i = 0; i < M::kLane; ++i) {
      T v[4];
      if (kl + i < n) {
        load_q4(xs + (kl + i) * kR, v);
      } else {
        v[0] = v[1] = v[2] = v[3] = (T)0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q][i] = v[q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) M::mma(acc[q], a, b[q]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = M::row_of(g, i);
    if (m >= kc) continue;
    T* __restrict__ zr = Z + (task.z_off + m) * kR;
    if constexpr (M::kZInterleaved) {
      store_q4(zr + 4 * c, (T)acc[0][i], (T)acc[1][i], (T)acc[2][i], (T)acc[3][i]);
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) zr[16 * q + c] = (T)acc[q][i];
    }
  }
}

// FP32 factors, one workgroup per leaf of at most 64 rows: the four
// wavefronts split the leaf's columns k in KS ranges and its rows in 4 / KS
// groups of KS row tiles, so every Z row the leaf needs is read by 4 / KS
// wavefronts (instead of four, one per 16-row tile); the KS partial tiles of
// a row group are then added through LDS in wavefront order, in FP64.
template <int KS>
__global__ __launch_bounds__(256) void hmm_leaf_ks_kernel(smx_hmatrix H, const float* __restrict__ Z,
                                                          double* __restrict__ YF) {
  using M = Mt<float>;
  constexpr int TR = KS;  // row tiles per wavefront
  __shared__ float red[KS > 1 ? 4 - 4 / KS : 1][TR * 16][64];  // partials of the non-leading wavefronts
  const int64_t L = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int kp = w % KS, rg = w / KS;  // column range, row group
  const int g = lane >> 4, c = lane & 15;
  const int32_t l0 = H.leaf_start[L], rows = H.leaf_stop[L] - l0;
  const int r0 = 16 * TR * rg;  // first row of this wavefront's group
  const int nt = rows > r0 ? ((rows - r0 + 15) >> 4 < TR ? (rows - r0 + 15) >> 4 : TR) : 0;
  const int64_t kw = H.leaf_kw[L];
  const int64_t z0 = H.leaf_zg_off[L];
  const int64_t K = H.leaf_zg_off[L + 1] - z0;
  const int32_t* __restrict__ zg = H.leaf_zg + z0;
  const int64_t ks = ((kw + KS * M::kStep - 1) / (KS * M::kStep)) * M::kStep;
  const int64_t kb = kp * ks;
  const int64_t ke = nt == 0 ? kb : (kb + ks < kw ? kb + ks : kw);
  const float* __restrict__ Ur[TR];
  bool row_ok[TR];
#pragma unroll
  for (int r = 0; r < TR; ++r) {
    const int row = r0 + 16 * r + c;
    row_ok[r] = row < rows;
    Ur[r] = reinterpret_cast<const float*>(H.U) + H.leaf_u_off[L] + (int64_t)(row_ok[r] ? row : 0) * kw;
  }
  M::acc_t acc[TR][4];
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[r][q][i] = 0.0f;
  for (int64_t kk = kb; kk < ke; kk += M::kStep) {
    const int64_t kc = kk + M::kLane * g;
    int64_t zi[M::kLane];
#pragma unroll
    for (int i = 0; i < M::kLane; ++i) zi[i] = (kc + i < K) ? (int64_t)zg[kc + i] : -1;
    float a[TR][M::kLane];
#pragma unroll
    for (int r = 0; r < TR; ++r) M::load_a(Ur[r] + kc, (row_ok[r] && r < nt) ? kw - kc : 0, a[r]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float b[M::kLane];
#pragma unroll
      for (int i = 0; i < M::kLane; ++i) b[i] = zi[i] >= 0 ? Z[zi[i] * kR + 16 * q + c] : 0.0f;
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        if (r >= nt) break;
        M::mma(acc[r][q], a[r], b);
      }
    }
  }
  if constexpr (KS > 1) {
    if (kp > 0) {
      const int slot = rg * (KS - 1) + kp - 1;
#pragma unroll
      for (int r = 0; r < TR; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int i = 0; i < 4; ++i) red[slot][16 * r + 4 * q + i][lane] = acc[r][q][i];
    }
    __syncthreads();
    if (kp > 0) return;
  }
#pragma unroll
  for (int r = 0; r < TR; ++r) {
    if (r >= nt) break;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int e = 16 * r + 4 * q + i;
        double v = (double)acc[r][q][i];
        if constexpr (KS > 1) {
#pragma unroll
          for (int j = 1; j < KS; ++j) v += (double)red[rg * (KS - 1) + j - 1][e][lane];
        }
        const int rr = r0 + 16 * r + M::row_of(g, i);
        if (rr < rows) YF[(int64_t)(l0 + rr) * kR + 16 * q + c] = v;
      }
  }
}

// Column ranges of the FP32 leaf launch (film, 64 right-hand sides, cache
// evicted): 1 (a 16-row tile per wavefront) 627 us, 4 (all rows per
// wavefront) 505 us; the one-tile hmm_leaf_kernel<float> took 595 us.
constexpr int kLeafKS = 4;

// One workgroup (four wavefronts) per leaf of at most 64 rows; wavefront w
// covers rows 16 w + (0..15).  The leaf's Z index list is staged in LDS,
// kLeafCols entries at a time, for all four wavefronts.
template <typename T>
__global__ __launch_bounds__(256) void hmm_leaf_kernel(smx_hmatrix H, const T* __restrict__ Z,
                                                       double* __restrict__ YF) {
  using M = Mt<T>;
  __shared__ int32_t zs[kLeafCols];
  const int64_t L = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int g = lane >> 4, c = lane & 15;
  const int32_t l0 = H.leaf_start[L], rows = H.leaf_stop[L] - l0;
  const int64_t kw = H.leaf_kw[L];
  const int64_t z0 = H.leaf_zg_off[L];
  const int64_t K = H.leaf_zg_off[L + 1] - z0;
  const int32_t* __restrict__ zg = H.leaf_zg + z0;
  const int r0 = 16 * w;
  const bool active = r0 < rows;  // wave-uniform; every wavefront takes the barriers
  const int row = r0 + c;
  const bool row_ok = row < rows;
  const T* __restrict__ Ur = reinterpret_cast<const T*>(H.U) + H.leaf_u_off[L] + (int64_t)(row_ok ? row : 0) * kw;
  typename M::acc_t acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[q][i] = 0;
  for (int64_t k0 = 0; k0 < kw; k0 += kLeafCols) {
    const int kn = (int)(K - k0 < kLeafCols ? (K - k0 > 0 ? K - k0 : 0) : kLeafCols);
    const int kspan = (int)(kw - k0 < kLeafCols ? kw - k0 : kLeafCols);
    __syncthreads();  // the previous chunk's indices are consumed
    for (int e = threadIdx.x; e < kn; e += 256) zs[e] = zg[k0 + e];
    __syncthreads();
    if (!active) continue;
    for (int kk = 0; kk < kspan; kk += M::kStep) {
      const int kc = kk + M::kLane * g;  // column within the chunk
      T a[M::kLane];
      M::load_a(Ur + k0 + kc, row_ok ? kw - k0 - kc : 0, a);
      int64_t zi[M::kLane];
#pragma unroll
      for (int i = 0; i < M::kLane; ++i) zi[i] = (kc + i < kn) ? (int64_t)zs[kc + i] : -1;
      if constexpr (!M::kZInterleaved) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          T b[M::kLane];
#pragma unroll
          for (int i = 0; i < M::kLane; ++i) b[i] = zi[i] >= 0 ? Z[zi[i] * kR + 16 * q + c] : (T)0;
          M::mma(acc[q], a, b);
        }
      } else {
        T b[4][M::kLane];
#pragma unroll
        for (int i = 0; i < M::kLane; ++i) {
          T v[4];
          if (zi[i] >= 0) {
            load_q4(Z + zi[i] * kR + 4 * c, v);
          } else {
            v[0] = v[1] = v[2] = v[3] = (T)0;
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) b[q][i] = v[q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) M::mma(acc[q], a, b[q]);
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rr = r0 + M::row_of(g, i);
      if (rr < rows) YF[(int64_t)(l0 + rr) * kR + 16 * q + c] = (double)acc[q][i];
    }
}

// One wavefront per permuted row, lane j = right-hand side j.
template <typename TN>
__global__ __launch_bounds__(256) void hmm_near_kernel(smx_hmatrix H, int32_t nrhs,
                                                       const double* __restrict__ XN,
                                                       const double* __restrict__ YF, double* __restrict__ Y,
                                                       int64_t ldy) {
  const int64_t p = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int j = threadIdx.x & 63;
  if (p >= H.nb) return;  // wave-uniform
  double s = H.diag_p[p] * XN[p * kR + j];
  if (H.near_ptr) {
    // the lanes load 64 entries of the row at once; each entry's value and
    // column then go to every lane by readlane (uniform index), so the
    // gathers of several entries are in flight together
    const TN* __restrict__ NV = reinterpret_cast<const TN*>(H.near_val);
    const int64_t e0 = H.near_ptr[p], e1 = H.near_ptr[p + 1];
    for (int64_t eb = e0; eb < e1; eb += 64) {
      const int64_t e = eb + j;
      const double v = e < e1 ? (double)NV[e] : 0.0;
      const int col = e < e1 ? H.near_idx[e] : 0;
      const int cnt = (int)((e1 - eb) < 64 ? (e1 - eb) : 64);
      const int lo = __double2loint(v), hi = __double2hiint(v);
#pragma unroll 4
      for (int t = 0; t < cnt; ++t) {
        const int ct = __builtin_amdgcn_readlane(col, t);
        const double vt = __hiloint2double(__builtin_amdgcn_readlane(hi, t), __builtin_amdgcn_readlane(lo, t));
        s = fma(vt, XN[(int64_t)ct * kR + j], s);
      }
    }
  }
  s += YF[p * kR + j];
  if (j < nrhs) Y[(int64_t)j * ldy + H.perm[p]] = s;
}

// Near field from dense tiles: one workgroup per tile, the tile's columns
// staged in LDS, its four wavefronts taking every fourth 16-column step (v_mfma_f64_16x16x4_f64 whatever the
// value and factor types; A = the tile's rows, row m of the tile on lane
// c = m; B = the staged X rows of its columns).  The four partial tiles meet
// in LDS and are summed in wavefront order, wavefront w finishing the
// right-hand sides 16 w .. 16 w + 15: diagonal and far sum added, the row
// written back over YF[p] (whole 128-byte segments); hmm_unstage_kernel then
// moves the rows to Y.
template <typename TN>
__device__ inline void load_near4(const TN* p, bool ok, double (&v)[4]);
template <>
__device__ inline void load_near4<double>(const double* p, bool ok, double (&v)[4]) {
  if (ok) {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    v[0] = a.x, v[1] = a.y, v[2] = b.x, v[3] = b.y;
  } else {
    v[0] = v[1] = v[2] = v[3] = 0.0;
  }
}
template <>
__device__ inline void load_near4<float>(const float* p, bool ok, double (&v)[4]) {
  if (ok) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
  } else {
    v[0] = v[1] = v[2] = v[3] = 0.0;
  }
}

template <typename TN>
__global__ __launch_bounds__(256) void hmm_near_tile_kernel(smx_hmatrix H, smx_hmat_near_tiles NT,
                                                            const double* __restrict__ XN,
                                                            double* __restrict__ YF) {
  using M = Mt<double>;
  __shared__ double part[4][4][4][64];  // [wavefront][q][i][lane], 32 KB
  __shared__ int32_t cs[kNtCols];       // the tile's columns, kNtCols at a time
  const int lane = threadIdx.x & 63;
  // tiles are stored in row order; XCD x (workgroups b with b % 8 == x) takes
  // the x-th eighth of them, so each XCD's L2 holds the X rows of one part of
  // the surface
  const int64_t t = xcd_item(NT.ntile);
  if (t >= NT.ntile) return;  // workgroup-uniform, before any barrier
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int g = lane >> 4, c = lane & 15;
  const smx_hmat_ntile tile = NT.tile[t];
  const int64_t w = tile.w;
  const int64_t ld = (w + 3) & ~(int64_t)3;
  const int nrows = tile.nrows;
  const bool row_ok = c < nrows;
  const TN* __restrict__ Ar = reinterpret_cast<const TN*>(NT.val) + tile.a_off + (int64_t)(row_ok ? c : 0) * ld;
  const int32_t* __restrict__ cl = NT.col + tile.col_off;
  const double* __restrict__ xs = XN + c;
  f64x4 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[q][i] = 0;
  for (int64_t k0 = 0; k0 < w; k0 += kNtCols) {
    const int kn = (int)(w - k0 < kNtCols ? w - k0 : kNtCols);
    __syncthreads();  // the previous chunk's columns are consumed
    for (int e = threadIdx.x; e < kn; e += 256) cs[e] = cl[k0 + e];
    __syncthreads();
    for (int kk = M::kStep * wv; kk < kn; kk += 4 * M::kStep) {
      const int kc = kk + M::kLane * g;  // column within the chunk
      double a[4];
      load_near4<TN>(Ar + k0 + kc, row_ok && k0 + kc < ld, a);
      int64_t ci[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) ci[i] = (kc + i < kn) ? (int64_t)cs[kc + i] : -1;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = ci[i] >= 0 ? xs[ci[i] * kR + 16 * q] : 0.0;
        M::mma(acc[q], a, b);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) part[wv][q][i][lane] = acc[q][i];
  __syncthreads();
  const int32_t* __restrict__ rows = NT.row + 16 * t;
  const int j = 16 * wv + c;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = M::row_of(g, i);
    if (m >= nrows) continue;
    const int64_t p = rows[m];
    double s = part[0][wv][i][lane];
    s += part[1][wv][i][lane];
    s += part[2][wv][i][lane];
    s += part[3][wv][i][lane];
    s += H.diag_p[p] * XN[p * kR + j] + YF[p * kR + j];
    YF[p * kR + j] = s;
  }
}

// Y[j][i] = YF[iperm[i]][j]: one workgroup per 64 output indices, the rows
// read whole (512 bytes each) into LDS and written out along i.
__global__ __launch_bounds__(256) void hmm_unstage_kernel(int64_t nb, int32_t nrhs, const double* __restrict__ YF,
                                                          const int32_t* __restrict__ iperm,
                                                          double* __restrict__ Y, int64_t ldy) {
  __shared__ double tb[64][kR + 1];
  const int64_t i0 = (int64_t)blockIdx.x * 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  for (int r = wv; r < 64; r += 4) {
    const int64_t i = i0 + r;
    if (i < nb) tb[r][lane] = YF[(int64_t)iperm[i] * kR + lane];
  }
  __syncthreads();
  const int64_t i = i0 + lane;
  if (i >= nb) return;
  for (int j = wv; j < nrhs; j += 4) Y[(int64_t)j * ldy + i] = tb[lane][j];
}

// Workspace: XR (factor type, nb x 64, columns interleaved), YF (FP64,
// nb x 64), Z (factor type, nz x 64, columns interleaved for FP64), XD (FP64 copy of
// X, nb x 64 in column order, for the near field and the diagonal), each
// 256-byte aligned.
struct MmLayout {
  int64_t xr, yf, z, xd;
};

MmLayout mm_layout(const smx_hmatrix* H, int64_t nz) {
  const int64_t ts = H->factor_dtype == SMX_F32 ? 4 : 8;
  auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
  MmLayout L;
  L.xr = up(ts * kR * H->nb);
  L.yf = up(8 * kR * H->nb);
  L.z = up(ts * kR * (nz > 0 ? nz : 1));
  L.xd = up(8 * kR * H->nb);  // FP64 copy of X in column order for the near field
  return L;
}

template <typename T>
int run_hmm(const smx_hmatrix* H, const smx_hmat_near_tiles* NT, int64_t nz, const double* X, int64_t ldx,
            int32_t nrhs, double* Y, int64_t ldy, char* work, hipStream_t st) {
  const MmLayout Lo = mm_layout(H, nz);
  T* XR = reinterpret_cast<T*>(work);
  double* YF = reinterpret_cast<double*>(work + Lo.xr);
  T* Z = reinterpret_cast<T*>(work + Lo.xr + Lo.yf);
  double* XD = reinterpret_cast<double*>(work + Lo.xr + Lo.yf + Lo.z);
  const double* XN = XD;
  hipLaunchKernelGGL(hmm_stage_kernel<T>, dim3((unsigned)((H->nb + 63) / 64)), dim3(256), 0, st, H->nb, nrhs, X,
                     ldx, H->gidx, XR, XD);
  SMX_LAUNCH_CHECK();
  if (H->nztask > 0 && nz > 0) {
    hipLaunchKernelGGL(hmm_z_kernel<T>, dim3((unsigned)((H->nztask + 3) / 4)), dim3(256), 0, st, *H, XR, Z);
    SMX_LAUNCH_CHECK();
  }
  if (H->nleaf > 0) {
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL(hmm_leaf_ks_kernel<kLeafKS>, dim3((unsigned)H->nleaf), dim3(256), 0, st, *H, Z, YF);
    else
      hipLaunchKernelGGL(hmm_leaf_kernel<T>, dim3((unsigned)H->nleaf), dim3(256), 0, st, *H, Z, YF);
    SMX_LAUNCH_CHECK();
  }
  if (NT) {
    if (NT->val_dtype == SMX_F32)
      hipLaunchKernelGGL(hmm_near_tile_kernel<float>, dim3(xcd_grid(NT->ntile)), dim3(256), 0, st, *H, *NT, XN, YF);
    else
      hipLaunchKernelGGL(hmm_near_tile_kernel<double>, dim3(xcd_grid(NT->ntile)), dim3(256), 0, st, *H, *NT, XN,
                         YF);
    SMX_LAUNCH_CHECK();
    hipLaunchKernelGGL(hmm_unstage_kernel, dim3((unsigned)((H->nb + 63) / 64)), dim3(256), 0, st, H->nb, nrhs, YF,
                       NT->iperm, Y, ldy);
    SMX_LAUNCH_CHECK();
    return SMX_OK;
  }
  const unsigned gn = (unsigned)((H->nb + 3) / 4);
  if (H->near_dtype == SMX_F32)
    hipLaunchKernelGGL(hmm_near_kernel<float>, dim3(gn), dim3(256), 0, st, *H, nrhs, XN, YF, Y, ldy);
  else
    hipLaunchKernelGGL(hmm_near_kernel<double>, dim3(gn), dim3(256), 0, st, *H, nrhs, XN, YF, Y, ldy);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

}  // namespace

extern "C" int64_t smx_hmat_matmat_ws_size(const smx_hmatrix* H, int64_t nz) {
  if (!H || H->nb <= 0) return 0;
  const MmLayout Lo = mm_layout(H, nz);
  return Lo.xr + Lo.yf + Lo.z + Lo.xd;
}

extern "C" int smx_hmat_matmat(const smx_hmatrix* H, const smx_hmat_near_tiles* NT, int64_t nz, const double* X,
                               int64_t ldx, int32_t nrhs, double* Y, int64_t ldy, void* work, int64_t work_bytes,
                               void* stream) {
  SMX_REQUIRE(H, SMX_ERR_VALUE, "null H");
  SMX_REQUIRE(H->nb >= 0 && H->nleaf >= 0 && H->nztask >= 0 && nz >= 0, SMX_ERR_SHAPE, "negative size");
  SMX_REQUIRE(nrhs >= 1 && nrhs <= kR, SMX_ERR_SHAPE, "nrhs %d must be in [1, %d]", nrhs, kR);
  if (H->nb == 0) return SMX_OK;
  SMX_REQUIRE(ldx >= H->nb && ldy >= H->nb, SMX_ERR_SHAPE, "ldx/ldy below nb");
  SMX_REQUIRE(X && Y && work && H->gidx && H->perm && H->diag_p && H->leaf_start && H->leaf_stop &&
                  H->leaf_kw && H->leaf_u_off && H->leaf_zg_off && H->U,
              SMX_ERR_VALUE, "null pointer argument");
  SMX_REQUIRE(H->factor_dtype == SMX_F32 || H->factor_dtype == SMX_F64, SMX_ERR_TYPE,
              "factor dtype must be FP32 or FP64");
  SMX_REQUIRE(H->near_dtype == SMX_F32 || H->near_dtype == SMX_F64, SMX_ERR_TYPE,
              "near dtype must be FP32 or FP64");
  SMX_REQUIRE(H->nztask == 0 || (H->ztask && H->V && H->leaf_zg), SMX_ERR_VALUE, "far-field arrays missing");
  SMX_REQUIRE(!H->near_ptr || (H->near_idx && H->near_val), SMX_ERR_VALUE, "near arrays missing");
  SMX_REQUIRE(H->nleaf <= 0x7ffffff0LL && H->nztask / 4 < 0x7ffffff0LL && H->nb * kR < 0x7fffffffLL * 256LL,
              SMX_ERR_SHAPE, "operator too large");
  if (NT) {
    SMX_REQUIRE(NT->ntile >= (H->nb + 15) / 16 && NT->ntile <= H->nb && NT->ntile <= 0x7ffffff0LL,
                SMX_ERR_SHAPE, "near tiles: %lld tiles for %lld rows", (long long)NT->ntile, (long long)H->nb);
    SMX_REQUIRE(NT->tile && NT->row && NT->col && NT->val && NT->iperm, SMX_ERR_VALUE, "near tile arrays missing");
    SMX_REQUIRE(NT->val_dtype == SMX_F32 || NT->val_dtype == SMX_F64, SMX_ERR_TYPE,
                "near tile dtype must be FP32 or FP64");
  }
  const int64_t need = smx_hmat_matmat_ws_size(H, nz);
  SMX_REQUIRE(work_bytes >= need, SMX_ERR_OOM, "workspace of %lld bytes, %lld needed", (long long)work_bytes,
              (long long)need);
  hipStream_t st = smx::as_stream(stream);
  if (H->factor_dtype == SMX_F32)
    return run_hmm<float>(H, NT, nz, X, ldx, nrhs, Y, ldy, reinterpret_cast<char*>(work), st);
  return run_hmm<double>(H, NT, nz, X, ldx, nrhs, Y, ldy, reinterpret_cast<char*>(work), st);
}
