// Lindholm double-layer assembly on gfx950.
//
// Weights follow Lindholm (1984), in the form of the reference's
// gpu/fields/Lindholm_kernel.py:30-125 (Van Oosterom-Strackee solid angle,
// log1p edge integrals, "opposite edge" weight convention).
//
// Mapping: one wavefront owns one target row r.  Its 64 lanes take 64
// different triangles of the relevant-triangle list; the (w0,w1,w2) of each
// (row, triangle) pair is accumulated into an LDS copy of the output row and
// the finished row is written with coalesced stores.  Lanes whose triangles
// share a vertex hit the same column, so the accumulator is 64-bit fixed
// point (units of 2^-60, range +-8) updated with integer LDS atomics:
// integer addition is associative, so every fill of a block gives the same
// bits whatever order the lanes land in.  Each weight is rounded once to
// 2^-60 (8.7e-19 absolute); the entries of B are below 1 in magnitude.  The
// triangle geometry (34 doubles per triangle) is staged once per chunk of 64
// triangles in LDS, structure-of-arrays, and shared by the 4 wavefronts
// (4 target rows) of a workgroup.  Everything is FP64: the weights are
// near-singular close to the target.
#include "smx_common.h"

namespace {

constexpr int kWaves = 4;
constexpr int kThreads = kWaves * smx::kWave;
constexpr int kChunk = 64;
constexpr int kTile = SMX_LINDHOLM_TILE_COLS;
constexpr double kTiny = 1e-300;
// fixed-point scale of the row accumulator: 2^60 units per 1.0
constexpr double kFixScale = 1152921504606846976.0;
constexpr double kFixInv = 1.0 / 1152921504606846976.0;
constexpr double kFixLimit = 8.0 * 1152921504606846976.0 - 4096.0;

__device__ __forceinline__ unsigned long long to_fixed(double w) {
  double q = w * kFixScale;
  q = q > kFixLimit ? kFixLimit : (q < -kFixLimit ? -kFixLimit : q);
  return (unsigned long long)__double2ll_rn(q);
}

__device__ __forceinline__ double p_log1p(double ri, double rj, double s) {
  double denom = ri + rj - s;
  if (denom < kTiny) denom = kTiny;
  return log1p((2.0 * s) / denom);
}

struct FillSmem {
  unsigned long long acc[kWaves][kTile];
  double g[SMX_TRI_FIELDS][kChunk];
  int32_t lcol[3][kChunk];
  int32_t lgid[3][kChunk];
};

// Rows [row0, row0+4) of one fill task; every thread of the workgroup calls
// this with the same task (it contains workgroup barriers).
__device__ __forceinline__ void fill_rows(
    FillSmem& sm, const double* __restrict__ Xb, const int32_t* __restrict__ rows, int64_t nrows,
    int64_t row0, const double* __restrict__ geom, int64_t ntot, const int32_t* __restrict__ tri_lid,
    const int32_t* __restrict__ rel_tri, const int32_t* __restrict__ rel_cols, int64_t nrel,
    int32_t ncols, double* __restrict__ out, int64_t ld_row, int64_t ld_col,
    int32_t incident_policy) {
  const int wid = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t r = row0 + wid;
  const bool active = r < nrows;

  for (int c = lane; c < ncols; c += smx::kWave) sm.acc[wid][c] = 0ull;

  int32_t grow = -1;
  double x0 = 0.0, x1 = 0.0, x2 = 0.0;
  if (active) {
    grow = rows[r];
    x0 = Xb[3 * (int64_t)grow + 0];
    x1 = Xb[3 * (int64_t)grow + 1];
    x2 = Xb[3 * (int64_t)grow + 2];
  }

  for (int64_t base = 0; base < nrel; base += kChunk) {
    const int nt = (int)((nrel - base) < kChunk ? (nrel - base) : kChunk);
    __syncthreads();  // previous chunk fully consumed
    for (int idx = threadIdx.x; idx < SMX_TRI_FIELDS * kChunk; idx += kThreads) {
      const int f = idx / kChunk;
      const int t = idx - f * kChunk;
      double v = 0.0;
      if (t < nt) {
        const int64_t gt = rel_tri ? (int64_t)rel_tri[base + t] : (base + t);
        v = geom[(int64_t)f * ntot + gt];
      }
      sm.g[f][t] = v;
    }
    for (int idx = threadIdx.x; idx < 3 * kChunk; idx += kThreads) {
      const int k = idx / kChunk;
      const int t = idx - k * kChunk;
      int32_t c = -1, gid = -1;
      if (t < nt) {
        const int64_t gt = rel_tri ? (int64_t)rel_tri[base + t] : (base + t);
        c = rel_cols[3 * (base + t) + k];
        gid = tri_lid[3 * gt + k];
      }
      sm.lcol[k][t] = c;
      sm.lgid[k][t] = gid;
    }
    __syncthreads();

    if (active && lane < nt) {
      const int t = lane;
      const double (&g)[SMX_TRI_FIELDS][kChunk] = sm.g;
      const bool skip = (incident_policy == 0) &&
                        (grow == sm.lgid[0][t] || grow == sm.lgid[1][t] || grow == sm.lgid[2][t]);
      if (!skip) {
        const double r0x = g[0][t] - x0, r0y = g[1][t] - x1, r0z = g[2][t] - x2;
        const double r1x = g[3][t] - x0, r1y = g[4][t] - x1, r1z = g[5][t] - x2;
        const double r2x = g[6][t] - x0, r2y = g[7][t] - x1, r2z = g[8][t] - x2;
        const double n0 = sqrt(r0x * r0x + r0y * r0y + r0z * r0z) + kTiny;
        const double n1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z) + kTiny;
        const double n2 = sqrt(r2x * r2x + r2y * r2y + r2z * r2z) + kTiny;
        const double h = g[9][t] * r0x + g[10][t] * r0y + g[11][t] * r0z;
        const double e0 = g[15][t] * r0x + g[16][t] * r0y + g[17][t] * r0z;
        const double e1 = g[18][t] * r1x + g[19][t] * r1y + g[20][t] * r1z;
        const double e2 = g[21][t] * r2x + g[22][t] * r2y + g[23][t] * r2z;
        const double s0 = g[12][t], s1 = g[13][t], s2 = g[14][t];
        const double P0 = p_log1p(n0, n1, s0);
        const double P1 = p_log1p(n1, n2, s1);
        const double P2 = p_log1p(n2, n0, s2);
        // Van Oosterom-Strackee solid angle
        const double det = r0x * (r1y * r2z - r1z * r2y) - r0y * (r1x * r2z - r1z * r2x) +
                           r0z * (r1x * r2y - r1y * r2x);
        const double d01 = r0x * r1x + r0y * r1y + r0z * r1z;
        const double d12 = r1x * r2x + r1y * r2y + r1z * r2z;
        const double d20 = r2x * r0x + r2y * r0y + r2z * r0z;
        const double den = n0 * n1 * n2 + d01 * n2 + d12 * n0 + d20 * n1;
        const double omega = 2.0 * atan2(det, den);
        const double gP0 = g[24][t] * P0 + g[25][t] * P1 + g[26][t] * P2;
        const double gP1 = g[27][t] * P0 + g[28][t] * P1 + g[29][t] * P2;
        const double gP2 = g[30][t] * P0 + g[31][t] * P1 + g[32][t] * P2;
        const double c = g[33][t];
        const double w0 = s1 * c * (e1 * omega - h * gP0);
        const double w1 = s2 * c * (e2 * omega - h * gP1);
        const double w2 = s0 * c * (e0 * omega - h * gP2);
        const int32_t c0 = sm.lcol[0][t], c1 = sm.lcol[1][t], c2 = sm.lcol[2][t];
        if (c0 >= 0) atomicAdd(&sm.acc[wid][c0], to_fixed(w0));
        if (c1 >= 0) atomicAdd(&sm.acc[wid][c1], to_fixed(w1));
        if (c2 >= 0) atomicAdd(&sm.acc[wid][c2], to_fixed(w2));
      }
    }
  }
  __syncthreads();
  if (active) {
    double* orow = out + r * ld_row;
    for (int c = lane; c < ncols; c += smx::kWave)
      orow[(int64_t)c * ld_col] = (double)(long long)sm.acc[wid][c] * kFixInv;
  }
}

__global__ __launch_bounds__(kThreads) void lindholm_fill_kernel(
    const double* __restrict__ Xb, const int32_t* __restrict__ rows, int64_t nrows,
    const double* __restrict__ geom, int64_t ntot, const int32_t* __restrict__ tri_lid,
    const int32_t* __restrict__ rel_tri, const int32_t* __restrict__ rel_cols, int64_t nrel,
    int32_t ncols, double* __restrict__ out, int64_t ld_row, int64_t ld_col,
    int32_t incident_policy) {
  __shared__ FillSmem sm;
  fill_rows(sm, Xb, rows, nrows, (int64_t)blockIdx.x * kWaves, geom, ntot, tri_lid, rel_tri,
            rel_cols, nrel, ncols, out, ld_row, ld_col, incident_policy);
}

__global__ __launch_bounds__(kThreads) void lindholm_fill_tasks_kernel(
    const double* __restrict__ Xb, const double* __restrict__ geom, int64_t ntot,
    const int32_t* __restrict__ tri_lid, const int32_t* __restrict__ rows,
    const int32_t* __restrict__ rel_tri, const int32_t* __restrict__ rel_cols,
    const smx_fill_task* __restrict__ tasks, const int32_t* __restrict__ wg_task,
    const int32_t* __restrict__ wg_row0, double* __restrict__ out, int32_t incident_policy) {
  __shared__ FillSmem sm;
  const smx_fill_task T = tasks[wg_task[blockIdx.x]];
  fill_rows(sm, Xb, rows + T.rows_off, T.nrows, wg_row0[blockIdx.x], geom, ntot, tri_lid,
            rel_tri + T.rel_off, rel_cols + 3 * T.rel_off, T.nrel, T.ncols, out + T.out_off,
            T.ld_row, T.ld_col, incident_policy);
}

__global__ void dense_finish_kernel(double* __restrict__ B, int64_t n, int64_t ld,
                                    const double* __restrict__ diag, void* __restrict__ lo,
                                    int32_t lo_dtype, int64_t r0) {
  const int64_t i = r0 + blockIdx.y;
  if (i >= n) return;
  double* row = B + i * ld;
  if (!lo) {
    if (diag && blockIdx.x == 0 && threadIdx.x == 0) row[i] = diag[i];
    return;
  }
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    const double v = (diag && j == i) ? diag[i] : row[j];
    if (diag && j == i) row[j] = v;
    if (lo_dtype == SMX_F32) {
      reinterpret_cast<float*>(lo)[i * ld + j] = (float)v;
    } else {
      reinterpret_cast<uint16_t*>(lo)[i * ld + j] = smx::f64_to_bf16_bits(v);
    }
  }
}

}  // namespace

extern "C" int smx_lindholm_fill(const double* Xb, const int32_t* rows, int64_t nrows,
                                 const double* tri_geom, int64_t ntri_total,
                                 const int32_t* tri_lid, const int32_t* rel_tri,
                                 const int32_t* rel_cols, int64_t nrel, int64_t ncols,
                                 double* out, int64_t ld_row, int64_t ld_col,
                                 int32_t incident_policy, void* stream) {
  SMX_REQUIRE(nrows >= 0 && nrel >= 0 && ncols >= 0, SMX_ERR_SHAPE, "negative size");
  SMX_REQUIRE(ncols <= kTile, SMX_ERR_SHAPE, "ncols %lld exceeds the tile width %d",
              (long long)ncols, kTile);
  SMX_REQUIRE(ntri_total > 0 || nrel == 0, SMX_ERR_SHAPE, "empty triangle table");
  if (nrows == 0 || ncols == 0) return SMX_OK;
  SMX_REQUIRE(Xb && rows && tri_geom && tri_lid && rel_cols && out, SMX_ERR_VALUE,
              "null pointer argument");
  const int grid = smx::grid_for(nrows, kWaves);
  hipLaunchKernelGGL(lindholm_fill_kernel, dim3(grid), dim3(kThreads), 0, smx::as_stream(stream),
                     Xb, rows, nrows, tri_geom, ntri_total, tri_lid, rel_tri, rel_cols, nrel,
                     (int32_t)ncols, out, ld_row, ld_col, incident_policy);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_lindholm_fill_batched(const double* Xb, const double* tri_geom,
                                         int64_t ntri_total, const int32_t* tri_lid,
                                         const int32_t* rows, const int32_t* rel_tri,
                                         const int32_t* rel_cols, const smx_fill_task* tasks,
                                         const int32_t* wg_task, const int32_t* wg_row0,
                                         int64_t nwg, double* out, int32_t incident_policy,
                                         void* stream) {
  SMX_REQUIRE(nwg >= 0, SMX_ERR_SHAPE, "negative workgroup count");
  if (nwg == 0) return SMX_OK;
  SMX_REQUIRE(nwg <= 0x7fffffffLL, SMX_ERR_SHAPE, "too many workgroups");
  SMX_REQUIRE(Xb && tri_geom && tri_lid && rows && rel_tri && rel_cols && tasks && wg_task &&
                  wg_row0 && out,
              SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(lindholm_fill_tasks_kernel, dim3((unsigned)nwg), dim3(kThreads), 0,
                     smx::as_stream(stream), Xb, tri_geom, ntri_total, tri_lid, rows, rel_tri,
                     rel_cols, tasks, wg_task, wg_row0, out, incident_policy);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_dense_finish(double* B, int64_t n, int64_t ld, const double* diag,
                                void* out_lo, int32_t lo_dtype, void* stream) {
  SMX_REQUIRE(n >= 0 && ld >= n, SMX_ERR_SHAPE, "bad shape n=%lld ld=%lld", (long long)n,
              (long long)ld);
  SMX_REQUIRE(out_lo == nullptr || lo_dtype == SMX_F32 || lo_dtype == SMX_BF16, SMX_ERR_TYPE,
              "lo_dtype must be FP32 or BF16");
  if (n == 0) return SMX_OK;
  int gx = smx::grid_for(n, 256);
  if (gx > 64) gx = 64;
  // grid.y is limited to 65535 rows per launch
  for (int64_t r0 = 0; r0 < n; r0 += 65535) {
    const int64_t nr = (n - r0) < 65535 ? (n - r0) : 65535;
    hipLaunchKernelGGL(dense_finish_kernel, dim3(gx, (unsigned)nr), dim3(256), 0,
                       smx::as_stream(stream), B, n, ld, diag, out_lo, lo_dtype, r0);
    SMX_LAUNCH_CHECK();
  }
  return SMX_OK;
}
