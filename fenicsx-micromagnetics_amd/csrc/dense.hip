// Dense boundary matvec y = B x on gfx950 (HBM-bound streaming GEMV).
//
// B is row-major with FP64, FP32 or BF16 storage.  A workgroup owns R rows;
// its wavefronts split the columns and stream them with 16-byte loads per lane
// (1 KiB per row per wavefront step, two steps in flight), so one x fragment
// read from L1/L2 is reused R times.  Products are
// converted to FP64 and accumulated in FP64, so FP32/BF16 storage only adds
// the storage rounding of B.  The final lane reduction is a butterfly on DPP /
// ds_swizzle shuffles.
#include "smx_common.h"

namespace {

constexpr int kWavesPerBlock = 4;

template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  static constexpr int N = 2;
};
template <>
struct Vec16<float> {
  static constexpr int N = 4;
};
template <>
struct Vec16<uint16_t> {
  static constexpr int N = 8;
};

__device__ __forceinline__ double to_f64(double v) { return v; }
__device__ __forceinline__ double to_f64(float v) { return (double)v; }
__device__ __forceinline__ double to_f64(uint16_t v) {
  return (double)__uint_as_float(((uint32_t)v) << 16);
}

struct alignas(16) Raw16 {
  uint32_t w[4];
};

template <typename T>
__device__ __forceinline__ void unpack16(const Raw16& r, double (&o)[Vec16<T>::N]);

template <>
__device__ __forceinline__ void unpack16<double>(const Raw16& r, double (&o)[2]) {
  o[0] = __hiloint2double((int)r.w[1], (int)r.w[0]);
  o[1] = __hiloint2double((int)r.w[3], (int)r.w[2]);
}
template <>
__device__ __forceinline__ void unpack16<float>(const Raw16& r, double (&o)[4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = (double)__uint_as_float(r.w[q]);
}
template <>
__device__ __forceinline__ void unpack16<uint16_t>(const Raw16& r, double (&o)[8]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    o[2 * q] = (double)__uint_as_float(r.w[q] << 16);
    o[2 * q + 1] = (double)__uint_as_float(r.w[q] & 0xffff0000u);
  }
}

// x staging: x[k] = src[gather[k]] - shift*scale (or src[k])
__global__ void gather_shift_kernel(int64_t n, const double* __restrict__ src,
                                    const int32_t* __restrict__ gather,
                                    const double* __restrict__ shift, double scale,
                                    double* __restrict__ xout) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double s = shift ? (*shift) * scale : 0.0;
  xout[k] = (gather ? src[gather[k]] : src[k]) - s;
}

// R rows per workgroup; the kWavesPerBlock wavefronts split the columns into
// contiguous quarters, so each workgroup streams R*n elements with every lane
// issuing R independent 16-byte loads per step (plus one shared x fragment),
// and n/R workgroups keep all CUs busy even for a few thousand rows.
// UN: 16-byte vectors per row and lane in flight per step (cube 150 nm,
// cache evicted: FP32 2 / 3 / 4 -> 27.6 / 33.8 / 28.9 us, FP64 44.5 / 44.1 /
// 44.5 us, BF16 20.8 / 23.5 / 18.4 us; the same bits)
template <typename T, int R, int UN>
__global__ __launch_bounds__(kWavesPerBlock * 64) void gemv_rows_kernel(
    const T* __restrict__ B, int64_t n, int64_t ldb, const double* __restrict__ x,
    double* __restrict__ y) {
  constexpr int V = Vec16<T>::N;
  __shared__ double part[kWavesPerBlock][R];
  const int wid = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t row0 = (int64_t)blockIdx.x * R;
  const int64_t nvec = n / V;  // full 16-byte vectors per row
  // each wavefront's column range starts on a 128-byte line (8 vectors)
  const int64_t per = ((nvec + 8 * kWavesPerBlock - 1) / (8 * kWavesPerBlock)) * 8;
  const int64_t v0 = wid * per;
  const int64_t v1 = (v0 + per < nvec) ? v0 + per : nvec;
  double acc[R];
#pragma unroll
  for (int q = 0; q < R; ++q) acc[q] = 0.0;
  const T* rowp[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int64_t rr = (row0 + q < n) ? row0 + q : n - 1;
    rowp[q] = B + rr * ldb;
  }
  int64_t v = v0 + lane;
  for (; v + 64 * (UN - 1) < v1; v += 64 * UN) {
    Raw16 raw[UN][R];
#pragma unroll
    for (int h = 0; h < UN; ++h)
#pragma unroll
      for (int q = 0; q < R; ++q) raw[h][q] = *reinterpret_cast<const Raw16*>(rowp[q] + (v + 64 * h) * V);
#pragma unroll
    for (int h = 0; h < UN; ++h) {
      double xv[V];
      const double* xp = x + (v + 64 * h) * V;
#pragma unroll
      for (int e = 0; e < V; ++e) xv[e] = xp[e];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        double bv[V];
        unpack16<T>(raw[h][q], bv);
#pragma unroll
        for (int e = 0; e < V; ++e) acc[q] = fma(bv[e], xv[e], acc[q]);
      }
    }
  }
  for (; v < v1; v += 64) {
    double xv[V];
    const double* xp = x + v * V;
#pragma unroll
    for (int e = 0; e < V; ++e) xv[e] = xp[e];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const Raw16 raw = *reinterpret_cast<const Raw16*>(rowp[q] + v * V);
      double bv[V];
      unpack16<T>(raw, bv);
#pragma unroll
      for (int e = 0; e < V; ++e) acc[q] = fma(bv[e], xv[e], acc[q]);
    }
  }
  if (wid == kWavesPerBlock - 1) {
    for (int64_t j = nvec * V + lane; j < n; j += 64) {
      const double xj = x[j];
#pragma unroll
      for (int q = 0; q < R; ++q) acc[q] = fma(to_f64(rowp[q][j]), xj, acc[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const double s = smx::wave_sum(acc[q]);
    if (lane == 0) part[wid][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < R && row0 + threadIdx.x < n) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) s += part[w][threadIdx.x];
    y[row0 + threadIdx.x] = s;
  }
}

// y = B^T x: block (bx, by) covers columns [64*bx*CV, ...) for rows chunk by
constexpr int kTRows = 256;
template <typename T>
__global__ __launch_bounds__(256) void gemv_t_partial_kernel(const T* __restrict__ B, int64_t n,
                                                             int64_t ldb,
                                                             const double* __restrict__ x,
                                                             double* __restrict__ work) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.y * kTRows;
  if (j >= n) return;
  const int64_t i1 = (i0 + kTRows < n) ? i0 + kTRows : n;
  double acc = 0.0;
  for (int64_t i = i0; i < i1; ++i) acc = fma(to_f64(B[i * ldb + j]), x[i], acc);
  work[(int64_t)blockIdx.y * n + j] = acc;
}

__global__ void gemv_t_reduce_kernel(int64_t n, int64_t nchunk, const double* __restrict__ work,
                                     double* __restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) s += work[c * n + j];
  y[j] = s;
}

template <typename T, int R, int UN = 2>
int launch_gemv(const void* B, int64_t n, int64_t ldb, const double* x, double* y,
                hipStream_t st) {
  constexpr int V = Vec16<T>::N;
  SMX_REQUIRE(ldb % V == 0, SMX_ERR_SHAPE, "ldb must be a multiple of %d for 16-byte rows", V);
  SMX_REQUIRE((reinterpret_cast<uintptr_t>(B) & 15) == 0, SMX_ERR_VALUE,
              "B must be 16-byte aligned");
  const int grid = smx::grid_for(n, R);
  hipLaunchKernelGGL((gemv_rows_kernel<T, R, UN>), dim3(grid), dim3(kWavesPerBlock * 64), 0, st,
                     reinterpret_cast<const T*>(B), n, ldb, x, y);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

}  // namespace

extern "C" int smx_dense_bmatvec(const void* B, int32_t dtype, int64_t n, int64_t ldb,
                                 const double* src, const int32_t* gather, const double* shift,
                                 double shift_scale, double* xbuf, double* y, void* stream) {
  SMX_REQUIRE(n >= 0 && ldb >= n, SMX_ERR_SHAPE, "bad shape n=%lld ldb=%lld", (long long)n,
              (long long)ldb);
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(B && src && y, SMX_ERR_VALUE, "null pointer argument");
  hipStream_t st = smx::as_stream(stream);
  const double* x = src;
  if (gather || shift) {
    SMX_REQUIRE(xbuf, SMX_ERR_VALUE, "xbuf required with gather/shift");
    hipLaunchKernelGGL(gather_shift_kernel, dim3(smx::grid_for(n, 256)), dim3(256), 0, st, n, src,
                       gather, shift, shift_scale, xbuf);
    SMX_LAUNCH_CHECK();
    x = xbuf;
  }
  switch (dtype) {
    case SMX_F64:
      return launch_gemv<double, 2>(B, n, ldb, x, y, st);
    case SMX_F32:
      // 4 rows per workgroup: 150 nm cube (5402 rows), cache evicted, 30.2 -> 28.0 us
      return launch_gemv<float, 4>(B, n, ldb, x, y, st);
    case SMX_BF16:
      return launch_gemv<uint16_t, 4, 4>(B, n, ldb, x, y, st);
    default:
      smx::set_error("unsupported dtype %d", dtype);
      return SMX_ERR_TYPE;
  }
}

extern "C" int smx_dense_bmatvec_t(const void* B, int32_t dtype, int64_t n, int64_t ldb,
                                   const double* x, double* y, double* work, void* stream) {
  SMX_REQUIRE(n >= 0 && ldb >= n, SMX_ERR_SHAPE, "bad shape");
  if (n == 0) return SMX_OK;
  SMX_REQUIRE(B && x && y && work, SMX_ERR_VALUE, "null pointer argument");
  hipStream_t st = smx::as_stream(stream);
  const int64_t nchunk = (n + kTRows - 1) / kTRows;
  SMX_REQUIRE(nchunk <= 65535, SMX_ERR_SHAPE, "n too large for the transposed matvec");
  dim3 grid(smx::grid_for(n, 256), (unsigned)nchunk);
  switch (dtype) {
    case SMX_F64:
      hipLaunchKernelGGL(gemv_t_partial_kernel<double>, grid, dim3(256), 0, st,
                         reinterpret_cast<const double*>(B), n, ldb, x, work);
      break;
    case SMX_F32:
      hipLaunchKernelGGL(gemv_t_partial_kernel<float>, grid, dim3(256), 0, st,
                         reinterpret_cast<const float*>(B), n, ldb, x, work);
      break;
    case SMX_BF16:
      hipLaunchKernelGGL(gemv_t_partial_kernel<uint16_t>, grid, dim3(256), 0, st,
                         reinterpret_cast<const uint16_t*>(B), n, ldb, x, work);
      break;
    default:
      smx::set_error("unsupported dtype %d", dtype);
      return SMX_ERR_TYPE;
  }
  SMX_LAUNCH_CHECK();
  hipLaunchKernelGGL(gemv_t_reduce_kernel, dim3(smx::grid_for(n, 256)), dim3(256), 0, st, n,
                     nchunk, work, y);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}
