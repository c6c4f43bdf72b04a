// SELL-64 sparse kernels (node-graph operators of the FEM discretisation).
//
// One thread per row; a slice of 64 rows is one wavefront, and entry j of the
// 64 rows of a slice is stored contiguously, so every index / value load of a
// wavefront is one coalesced 256 B (index) or 64*8*k B (values) request.
#include "smx_common.h"

namespace {

constexpr int kBlock = 256;

template <int KIND>
__global__ __launch_bounds__(kBlock) void sell_spmv_kernel(
    int64_t nrows, const int64_t* __restrict__ sp, const int32_t* __restrict__ cols,
    const double* __restrict__ vals, const double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ partials) {
  __shared__ double scratch[kBlock / 64];
  const int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t nslice = (nrows + 63) >> 6;
  double total = 0.0;
  if (row < nslice * 64) {
    const int64_t s = row >> 6;
    const int lane = (int)(row & 63);
    const int64_t start = sp[s];
    const int64_t w = (sp[s + 1] - start) >> 6;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t j = 0; j < w; ++j) {
      const int64_t e = start + j * 64 + lane;
      const int64_t c = cols[e];
      if (KIND == 0) {
        a0 = fma(vals[e], x[c], a0);
      } else if (KIND == 1) {
        a0 = fma(vals[3 * e + 0], x[3 * c + 0], a0);
        a0 = fma(vals[3 * e + 1], x[3 * c + 1], a0);
        a0 = fma(vals[3 * e + 2], x[3 * c + 2], a0);
      } else if (KIND == 2) {
        const double xc = x[c];
        a0 = fma(vals[3 * e + 0], xc, a0);
        a1 = fma(vals[3 * e + 1], xc, a1);
        a2 = fma(vals[3 * e + 2], xc, a2);
      } else {
        const double* A = vals + 9 * e;
        const double m0 = x[3 * c], m1 = x[3 * c + 1], m2 = x[3 * c + 2];
        a0 += A[0] * m0 + A[1] * m1 + A[2] * m2;
        a1 += A[3] * m0 + A[4] * m1 + A[5] * m2;
        a2 += A[6] * m0 + A[7] * m1 + A[8] * m2;
      }
    }
    if (row < nrows) {
      if (KIND <= 1) {
        y[row] = a0;
        total = a0;
      } else {
        y[3 * row + 0] = a0;
        y[3 * row + 1] = a1;
        y[3 * row + 2] = a2;
      }
    }
  }
  if (KIND <= 1 && partials) {
    double v[1] = {total};
    smx::block_sum<1>(v, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
  }
}

}  // namespace

extern "C" int64_t smx_reduce_ws_size(int64_t n) {
  return smx::red_ws_doubles((n + 63) / 64 + 1, 8);
}

extern "C" int64_t smx_sell_spmv_blocks(int64_t nrows) {
  const int64_t nslice = (nrows + 63) / 64;
  return (nslice * 64 + kBlock - 1) / kBlock;
}

extern "C" int smx_sell_spmv(int64_t nrows, const int64_t* slice_ptr, const int32_t* cols,
                             const double* vals, int32_t kind, const double* x, double* y,
                             double* partials, void* stream) {
  SMX_REQUIRE(nrows >= 0, SMX_ERR_SHAPE, "negative size");
  SMX_REQUIRE(kind >= 0 && kind <= 3, SMX_ERR_VALUE, "bad kind %d", kind);
  SMX_REQUIRE(!partials || kind <= 1, SMX_ERR_VALUE, "partial sums need kind 0 or 1");
  if (nrows == 0) return SMX_OK;
  SMX_REQUIRE(slice_ptr && cols && vals && x && y, SMX_ERR_VALUE, "null pointer argument");
  hipStream_t st = smx::as_stream(stream);
  const int64_t nslice = (nrows + 63) / 64;
  const int grid = smx::grid_for(nslice * 64, kBlock);
  switch (kind) {
    case 0:
      hipLaunchKernelGGL(sell_spmv_kernel<0>, dim3(grid), dim3(kBlock), 0, st, nrows, slice_ptr,
                         cols, vals, x, y, partials);
      break;
    case 1:
      hipLaunchKernelGGL(sell_spmv_kernel<1>, dim3(grid), dim3(kBlock), 0, st, nrows, slice_ptr,
                         cols, vals, x, y, partials);
      break;
    case 2:
      hipLaunchKernelGGL(sell_spmv_kernel<2>, dim3(grid), dim3(kBlock), 0, st, nrows, slice_ptr,
                         cols, vals, x, y, nullptr);
      break;
    default:
      hipLaunchKernelGGL(sell_spmv_kernel<3>, dim3(grid), dim3(kBlock), 0, st, nrows, slice_ptr,
                         cols, vals, x, y, nullptr);
      break;
  }
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}
