// Shared helpers for the gfx950 kernels of libsmx_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/smx_hip.h"

namespace smx {

constexpr int kWave = 64;

// Last error message of the calling host thread (read with smx_last_error()).
void set_error(const char* fmt, ...);

#define SMX_HIP_CHECK(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::smx::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                       __FILE__, __LINE__);                                   \
      return SMX_ERR_HIP;                                                     \
    }                                                                         \
  } while (0)

#define SMX_LAUNCH_CHECK()                                                    \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess) {                                                   \
      ::smx::set_error("kernel launch failed: %s (%s:%d)",                    \
                       hipGetErrorString(_e), __FILE__, __LINE__);            \
      return SMX_ERR_HIP;                                                     \
    }                                                                         \
  } while (0)

#define SMX_REQUIRE(cond, code, ...)                                          \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::smx::set_error(__VA_ARGS__);                                          \
      return (code);                                                          \
    }                                                                         \
  } while (0)

// FP64 -> BF16 with a single round-to-nearest-even: round to FP32 with
// round-to-odd (truncate, then set the last bit when inexact), then round that
// FP32 to BF16 (Boldo & Melquiond, "Emulation of FMA and correctly rounded
// sums: proved algorithms using rounding to odd", 2008: with at least two
// extra bits the second rounding is the correct one).  Returns the bits.
__device__ __forceinline__ uint16_t f64_to_bf16_bits(double v) {
  float f = (float)v;
  if ((double)f != v && f == f && (__float_as_uint(f) & 0x7f800000u) != 0x7f800000u) {
    if (fabs((double)f) > fabs(v)) f = __uint_as_float(__float_as_uint(f) - 1u);  // toward zero
    f = __uint_as_float(__float_as_uint(f) | 1u);                                  // odd
  }
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Wavefront total through DPP row operations (VALU lane moves, no LDS round
// trip): pairs, quads, 8-lane halves and 16-lane rows by quad_perm and the row
// mirrors (each step adds a lane that holds an equal-sized disjoint partial),
// then the four row totals by readlane in a fixed order.  The result is
// uniform and identical in every lane.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

__device__ __forceinline__ double wave_total(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x140>(v);  // row_mirror
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Error-free sum (Knuth's TwoSum): s + e == a + b exactly, s = fl(a + b).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bv = s - a;
  e = (a - (s - bv)) + (b - bv);
}

// Compensated accumulation of v into (s, c): s + c carries the running sum
// with the rounding error of every addition kept in c (Neumaier's variant).
__device__ __forceinline__ void comp_add(double& s, double& c, double v) {
  double t, e;
  two_sum(s, v, t, e);
  s = t;
  c += e;
}

// Workgroup barrier that orders LDS accesses only: outstanding global loads
// and stores of the calling wave are not waited for.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Deterministic block sum of NV doubles per thread; result valid in thread 0.
// `scratch` must hold NV * (blockDim.x / 64) doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* scratch) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) scratch[q * nw + wid] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += scratch[q * nw + w];
      v[q] = s;
    }
  }
  __syncthreads();
}

// Grid-wide deterministic reduction ("last block" pattern).  `red` layout:
// red[0] reinterpreted as the arrival counter (must be 0 on entry; the last
// block resets it), partials of quantity q of block b at red[8 + q*G + b].
// Every block calls this once with its per-thread values; it returns true in
// the block that arrives last, whose thread 0 then holds the grid totals in
// `v`.  Totals are summed in a fixed order independent of arrival order.
template <int NV>
__device__ __forceinline__ bool grid_sum_last(double (&v)[NV], double* red, double* scratch) {
  __shared__ bool s_last;
  block_sum<NV>(v, scratch);
  const int G = gridDim.x;
  double* part = red + 8;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) part[q * G + blockIdx.x] = v[q];
    __threadfence();
    const unsigned t = atomicAdd(reinterpret_cast<unsigned*>(red), 1u);
    s_last = (t == (unsigned)(G - 1));
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < G; b += blockDim.x) {
      s += __hip_atomic_load(part + q * G + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    v[q] = s;
  }
  block_sum<NV>(v, scratch);
  if (threadIdx.x == 0) {
    __hip_atomic_store(reinterpret_cast<unsigned*>(red), 0u, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

// Doubles of reduction workspace needed for NV quantities over G blocks.
inline int64_t red_ws_doubles(int64_t G, int NV) { return 8 + (int64_t)NV * G; }

inline int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (int)(g < 1 ? 1 : g);
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace smx
