// Multi-RHS dense boundary product Y = B X on the matrix cores.
//
// B is the dense Lindholm matrix (n x n row-major, leading dimension ldb a
// multiple of 8, zero padded), X and Y hold nrhs <= 64 right-hand sides as
// columns (column j of X at X + j*ldx).  One workgroup (4 wavefronts) owns a
// 64-row x 64-RHS output tile over one K slice; each wavefront owns 16 rows
// and four 16 x 16 accumulator tiles, the X slice of a step is shared
// through LDS.  Split-K partial sums go to a workspace
// and are added in slice order by a second kernel, so the result does not
// depend on scheduling.
//
// MFMA per storage type (lane l, g = l / 16):
//   BF16: v_mfma_f32_16x16x32_bf16, A[m = l%16][k], B[k][n = l%16] with
//         k = 8g + i (i < 8), D[m = 4g + i][n = l%16] (FP32 accumulate)
//   FP32: v_mfma_f32_16x16x4_f32, one k per lane (k = g), same D layout
//   FP64: v_mfma_f64_16x16x4_f64, one k per lane (k = g),
//         D[m = g + 4i][n = l%16] (FP64 accumulate)
// The K order inside a step is permuted so that every lane streams 32
// contiguous bytes of its B row and of each X column per step: MFMA j of a
// step uses k = kk + KL*g + (sub-range j), identical for both operands.
#include <hip/hip_bf16.h>

#include "smx_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 64;  // rows per workgroup tile
constexpr int kRhs = 64;   // right-hand sides per launch

template <typename T>
struct Mm;

// BF16: step of 64 k; each lane loads 16 consecutive k (32 B) = 2 MFMAs of 8.
template <>
struct Mm<__bf16> {
  using acc_t = f32x4;
  static constexpr int kStep = 64;
  static constexpr int kLane = 16;  // k per lane per step
  struct Frag {
    bf16x8 v[2];
  };
  // `room` = entries left in the row from p; each 8-entry half (16 B) is
  // loaded only if it lies inside (room is a multiple of 8)
  __device__ static Frag load(const __bf16* p, int64_t room) {
    Frag f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (8 * h < room) {
        f.v[h] = *reinterpret_cast<const bf16x8*>(p + 8 * h);
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) f.v[h][i] = (__bf16)0.0f;
      }
    }
    return f;
  }
  __device__ static void mma(acc_t& c, const Frag& a, const Frag& b) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v[0], b.v[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v[1], b.v[1], c, 0, 0, 0);
  }
  __device__ static int row_of(int g, int i) { return 4 * g + i; }
};

// FP32: step of 32 k; each lane loads 8 consecutive k = 8 MFMAs of 4.
template <>
struct Mm<float> {
  using acc_t = f32x4;
  static constexpr int kStep = 32;
  static constexpr int kLane = 8;
  struct Frag {
    float v[8];
  };
  __device__ static Frag load(const float* p, int64_t room) {  // room % 8 == 0: all or nothing
    Frag f;
    if (room > 0) {
      const float4 a = *reinterpret_cast<const float4*>(p);
      const float4 b = *reinterpret_cast<const float4*>(p + 4);
      f.v[0] = a.x, f.v[1] = a.y, f.v[2] = a.z, f.v[3] = a.w;
      f.v[4] = b.x, f.v[5] = b.y, f.v[6] = b.z, f.v[7] = b.w;
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) f.v[i] = 0.0f;
    }
    return f;
  }
  __device__ static void mma(acc_t& c, const Frag& a, const Frag& b) {
#pragma unroll
    for (int j = 0; j < 8; ++j) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v[j], b.v[j], c, 0, 0, 0);
  }
  __device__ static int row_of(int g, int i) { return 4 * g + i; }
};

// FP64: step of 16 k; each lane loads 4 consecutive k = 4 MFMAs of 4.
template <>
struct Mm<double> {
  using acc_t = f64x4;
  static constexpr int kStep = 16;
  static constexpr int kLane = 4;
  struct Frag {
    double v[4];
  };
  __device__ static Frag load(const double* p, int64_t room) {  // room % 4 == 0: all or nothing
    Frag f;
    if (room > 0) {
      const double2 a = *reinterpret_cast<const double2*>(p);
      const double2 b = *reinterpret_cast<const double2*>(p + 2);
      f.v[0] = a.x, f.v[1] = a.y, f.v[2] = b.x, f.v[3] = b.y;
    } else {
      f.v[0] = f.v[1] = f.v[2] = f.v[3] = 0.0;
    }
    return f;
  }
  __device__ static void mma(acc_t& c, const Frag& a, const Frag& b) {
#pragma unroll
    for (int j = 0; j < 4; ++j) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a.v[j], b.v[j], c, 0, 0, 0);
  }
  __device__ static int row_of(int g, int i) { return g + 4 * i; }
};

template <typename T>
using acc_elem_t = typename std::conditional<std::is_same<T, double>::value, double, float>::type;

// Xt[j][k] = X[j*ldx + k] in T for k < n, 0 for n <= k < kpad, 0 for j >= nrhs.
template <typename T>
__global__ void mm_stage_x_kernel(int64_t n, int64_t kpad, int32_t nrhs, const double* __restrict__ X,
                                  int64_t ldx, T* __restrict__ Xt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kpad * kRhs) return;
  const int64_t j = i / kpad, k = i - j * kpad;
  const double v = (j < nrhs && k < n) ? X[j * ldx + k] : 0.0;
  if constexpr (std::is_same<T, __bf16>::value) {
    reinterpret_cast<uint16_t*>(Xt)[i] = smx::f64_to_bf16_bits(v);  // one rounding, as B's
  } else {
    Xt[i] = (T)v;
  }
}

// part[s][row][rhs] = sum over the K slice s of B[row][k] Xt[rhs][k].  Each
// step's X slice (64 right-hand sides x kStep k) is staged in LDS once per
// workgroup (double-buffered, one barrier per step) and shared by the four
// wavefronts; every wavefront streams its own 16 rows of B.  Rows of the
// staged slice are padded by 16 bytes against bank conflicts.  Against every
// wavefront reading the slice itself from L2 (cube 150 nm, 64 right-hand
// sides, whole product): FP32 87.8 -> 62.1 us, FP64 170.1 -> 110.9 us, BF16
// 53.8 -> 36.3 us (the library GEMM of the same shape: 57.2, 193.8, 33.4 us).
template <typename T>
__global__ __launch_bounds__(256) void mm_tile_lds_kernel(int64_t n, const T* __restrict__ B, int64_t ldb,
                                                          const T* __restrict__ Xt, int64_t kpad,
                                                          int64_t kslice, int64_t npad,
                                                          acc_elem_t<T>* __restrict__ part) {
  using M = Mm<T>;
  constexpr int kRowElems = M::kStep + 16 / (int)sizeof(T);  // padded LDS row (elements)
  __shared__ __align__(16) T xs[2][kRhs * kRowElems];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t row0 = (int64_t)blockIdx.x * kRows + 16 * w;
  const int s = blockIdx.y;
  const int64_t k0 = s * kslice;
  const int64_t k1 = (k0 + kslice) < kpad ? (k0 + kslice) : kpad;
  int64_t r = row0 + c;
  if (r > n - 1) r = n - 1;  // rows past n load row n-1; their sums are not stored
  const T* __restrict__ Brow = B + r * ldb;
  typename M::acc_t acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[t][i] = 0;
  // staging: each of the 64 slice rows holds kStep * sizeof(T) = 128 bytes
  // (every type), eight 16-byte pieces: thread i takes row i / 4, pieces
  // 2 (i % 4) and 2 (i % 4) + 1.  The next slice is loaded into registers at
  // the top of a step and written to the other buffer after its MFMAs.
  constexpr int kPieceElems = 16 / (int)sizeof(T);
  const int sr = threadIdx.x >> 2, sp = 2 * (threadIdx.x & 3);
  const T* __restrict__ xsrc = Xt + (int64_t)sr * kpad + sp * kPieceElems;
  T* const xdst0 = &xs[0][sr * kRowElems + sp * kPieceElems];
  T* const xdst1 = &xs[1][sr * kRowElems + sp * kPieceElems];
  int buf = 0;
  if (k0 < k1) {
    const float4 v0 = *reinterpret_cast<const float4*>(xsrc + k0);
    const float4 v1 = *reinterpret_cast<const float4*>(xsrc + k0 + kPieceElems);
    *reinterpret_cast<float4*>(xdst0) = v0;
    *reinterpret_cast<float4*>(xdst0 + kPieceElems) = v1;
  }
  __syncthreads();
  for (int64_t kk = k0; kk < k1; kk += M::kStep) {
    const int64_t kl = kk + M::kLane * g;
    const typename M::Frag a = M::load(Brow + kl, ldb - kl);
    const bool more = kk + M::kStep < k1;
    float4 n0 = make_float4(0.f, 0.f, 0.f, 0.f), n1 = n0;
    if (more) {
      n0 = *reinterpret_cast<const float4*>(xsrc + kk + M::kStep);
      n1 = *reinterpret_cast<const float4*>(xsrc + kk + M::kStep + kPieceElems);
    }
    typename M::Frag x[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const T* p = &xs[buf][(16 * t + c) * kRowElems + M::kLane * g];
      x[t] = M::load(p, M::kLane);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) M::mma(acc[t], a, x[t]);
    if (more) {  // the other buffer was last read before the previous barrier
      T* d = buf ? xdst0 : xdst1;
      *reinterpret_cast<float4*>(d) = n0;
      *reinterpret_cast<float4*>(d + kPieceElems) = n1;
    }
    __syncthreads();
    buf ^= 1;
  }
  acc_elem_t<T>* __restrict__ P = part + (int64_t)s * npad * kRhs;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t row = row0 + M::row_of(g, i);
      P[row * kRhs + 16 * t + c] = acc[t][i];
    }
}

// Y[j*ldy + row] = sum_s part[s][row][j] (slice order, FP64)
template <typename A>
__global__ void mm_reduce_kernel(int64_t n, int32_t nrhs, int32_t nslice, int64_t npad,
                                 const A* __restrict__ part, double* __restrict__ Y, int64_t ldy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kRhs) return;
  const int64_t row = i / kRhs;
  const int j = (int)(i - row * kRhs);
  if (j >= nrhs) return;
  double s = 0.0;
  for (int q = 0; q < nslice; ++q) s += (double)part[((int64_t)q * npad + row) * kRhs + j];
  Y[(int64_t)j * ldy + row] = s;
}

struct MmPlan {
  int64_t kpad, npad, kslice, nslice, x_bytes, part_bytes;
};

template <typename T>
MmPlan plan(int64_t n) {
  using M = Mm<T>;
  MmPlan p;
  p.kpad = (n + 63) / 64 * 64;
  p.npad = (n + kRows - 1) / kRows * kRows;
  const int64_t tiles = p.npad / kRows;
  // about four workgroups per CU (cube FP32: 512, 768, 1024, 2048 workgroups
  // 72.5, 67.3, 62.1, 63.6 us)
  int64_t want = (1024 + tiles - 1) / tiles;
  if (want < 1) want = 1;
  if (want > 16) want = 16;
  int64_t ks = (p.kpad + want - 1) / want;
  ks = (ks + M::kStep - 1) / M::kStep * M::kStep;
  if (ks < M::kStep) ks = M::kStep;
  p.kslice = ks;
  p.nslice = (p.kpad + ks - 1) / ks;
  p.x_bytes = (int64_t)sizeof(T) * p.kpad * kRhs;
  p.x_bytes = (p.x_bytes + 255) / 256 * 256;
  p.part_bytes = (int64_t)sizeof(acc_elem_t<T>) * p.nslice * p.npad * kRhs;
  return p;
}

template <typename T>
int run_mm(const void* Bv, int64_t n, int64_t ldb, const double* X, int64_t ldx, int32_t nrhs, double* Y,
           int64_t ldy, void* work, int64_t work_bytes, hipStream_t st) {
  const MmPlan p = plan<T>(n);
  SMX_REQUIRE(work_bytes >= p.x_bytes + p.part_bytes, SMX_ERR_OOM,
              "workspace of %lld bytes, %lld needed", (long long)work_bytes,
              (long long)(p.x_bytes + p.part_bytes));
  T* Xt = reinterpret_cast<T*>(work);
  auto* part = reinterpret_cast<acc_elem_t<T>*>(reinterpret_cast<char*>(work) + p.x_bytes);
  const int64_t nx = p.kpad * kRhs;
  hipLaunchKernelGGL(mm_stage_x_kernel<T>, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, st, n,
                     p.kpad, nrhs, X, ldx, Xt);
  SMX_LAUNCH_CHECK();
  hipLaunchKernelGGL(mm_tile_lds_kernel<T>, dim3((unsigned)(p.npad / kRows), (unsigned)p.nslice), dim3(256),
                     0, st, n, reinterpret_cast<const T*>(Bv), ldb, Xt, p.kpad, p.kslice, p.npad, part);
  SMX_LAUNCH_CHECK();
  const int64_t ny = n * kRhs;
  hipLaunchKernelGGL(mm_reduce_kernel<acc_elem_t<T>>, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0,
                     st, n, nrhs, (int32_t)p.nslice, p.npad, part, Y, ldy);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

}  // namespace

extern "C" int64_t smx_dense_bmatmul_ws_size(int64_t n, int32_t dtype) {
  if (n <= 0) return 0;
  MmPlan p;
  if (dtype == SMX_BF16)
    p = plan<__bf16>(n);
  else if (dtype == SMX_F32)
    p = plan<float>(n);
  else
    p = plan<double>(n);
  return p.x_bytes + p.part_bytes;
}

extern "C" int smx_dense_bmatmul(const void* B, int32_t dtype, int64_t n, int64_t ldb, const double* X,
                                 int64_t ldx, int32_t nrhs, double* Y, int64_t ldy, void* work,
                                 int64_t work_bytes, void* stream) {
  SMX_REQUIRE(n >= 0 && ldb >= n && ldb % 8 == 0, SMX_ERR_SHAPE,
              "need ldb >= n and ldb %% 8 == 0 (n=%lld ldb=%lld)", (long long)n, (long long)ldb);
  SMX_REQUIRE(nrhs >= 1 && nrhs <= kRhs, SMX_ERR_SHAPE, "nrhs must be in [1, %d]", kRhs);
  SMX_REQUIRE(ldx >= n && ldy >= n, SMX_ERR_SHAPE, "ldx and ldy must be >= n");
  SMX_REQUIRE(dtype == SMX_F64 || dtype == SMX_F32 || dtype == SMX_BF16, SMX_ERR_TYPE, "bad dtype");
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(B && X && Y && work, SMX_ERR_VALUE, "null pointer argument");
  SMX_REQUIRE(n < (1LL << 31), SMX_ERR_SHAPE, "n too large");
  hipStream_t st = smx::as_stream(stream);
  if (dtype == SMX_BF16) return run_mm<__bf16>(B, n, ldb, X, ldx, nrhs, Y, ldy, work, work_bytes, st);
  if (dtype == SMX_F32) return run_mm<float>(B, n, ldb, X, ldx, nrhs, Y, ldy, work, work_bytes, st);
  return run_mm<double>(B, n, ldb, X, ldx, nrhs, Y, ldy, work, work_bytes, st);
}
