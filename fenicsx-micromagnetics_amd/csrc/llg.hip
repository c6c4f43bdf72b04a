// Fused effective field + LLG right-hand side, Jacobian-vector product,
// energy reductions and the Dormand-Prince stage combination.
//
// The field kernel walks the node graph once per node: exchange (k=1 values),
// DMI (3x3 blocks) and the demag gradient G (k=3) share one SELL-64 pattern,
// so a single pass gathers m_j and phi_j = u1_j + u2_j of the neighbours and
// produces H_eff and the LLG RHS without writing the intermediate fields.
#include <cstdarg>

#include "smx_common.h"

namespace smx {
static thread_local char g_err[512] = "";
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace smx

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void cross3(double ax, double ay, double az, double bx, double by,
                                       double bz, double& cx, double& cy, double& cz) {
  cx = ay * bz - az * by;
  cy = az * bx - ax * bz;
  cz = ax * by - ay * bx;
}

// Cubic axes of node k (9 doubles: c1, c2, c3).
__device__ __forceinline__ void cubic_axes(const smx_field_args& a, int64_t k, double (&c)[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) c[i] = a.cub_axes ? a.cub_axes[9 * k + i] : a.cub_axes_u[i];
}

// H_cub = pref (s1 c1 + s2 c2 + s3 c3), a_i = m.c_i, s1 = a1 (a2^2 + a3^2)
// (cyclic); with v != NULL the derivative dH_cub/dm v instead.
__device__ __forceinline__ void cubic_term(const smx_field_args& a, int64_t k, double mx, double my,
                                           double mz, const double* v, double& hx, double& hy,
                                           double& hz) {
  double c[9];
  cubic_axes(a, k, c);
  const double a1 = mx * c[0] + my * c[1] + mz * c[2];
  const double a2 = mx * c[3] + my * c[4] + mz * c[5];
  const double a3 = mx * c[6] + my * c[7] + mz * c[8];
  double s1, s2, s3;
  if (!v) {
    s1 = a1 * (a2 * a2 + a3 * a3);
    s2 = a2 * (a3 * a3 + a1 * a1);
    s3 = a3 * (a1 * a1 + a2 * a2);
  } else {
    const double vx = v[0], vy = v[1], vz = v[2];
    const double d1 = vx * c[0] + vy * c[1] + vz * c[2];
    const double d2 = vx * c[3] + vy * c[4] + vz * c[5];
    const double d3 = vx * c[6] + vy * c[7] + vz * c[8];
    s1 = d1 * (a2 * a2 + a3 * a3) + 2.0 * a1 * (a2 * d2 + a3 * d3);
    s2 = d2 * (a3 * a3 + a1 * a1) + 2.0 * a2 * (a3 * d3 + a1 * d1);
    s3 = d3 * (a1 * a1 + a2 * a2) + 2.0 * a3 * (a1 * d1 + a2 * d2);
  }
  const double p = a.cub_pref;
  hx += p * (s1 * c[0] + s2 * c[3] + s3 * c[6]);
  hy += p * (s1 * c[1] + s2 * c[4] + s3 * c[7]);
  hz += p * (s1 * c[2] + s2 * c[5] + s3 * c[8]);
}

__device__ __forceinline__ void sot_p(const smx_field_args& a, int64_t k, double& px, double& py,
                                      double& pz) {
  if (a.sot_p) {
    px = a.sot_p[3 * k];
    py = a.sot_p[3 * k + 1];
    pz = a.sot_p[3 * k + 2];
  } else {
    px = a.sot_p_u[0];
    py = a.sot_p_u[1];
    pz = a.sot_p_u[2];
  }
}

// Row i of the DMI block at graph slot e times (v0, v1, v2): the components
// present in the layout, in column order, the first as a product and the
// rest by fma (absent components are exact zeros), so the full and the
// compact layouts give the same bits.  0 when the row has no component.
__device__ __forceinline__ double dmi_row(const smx_field_args& a, int64_t e, int i, double v0, double v1,
                                          double v2) {
#pragma clang fp contract(off)
  const int nc = a.dmi_ncomp > 0 ? a.dmi_ncomp : 9;
  const double* __restrict__ D = a.dmi_vals + (int64_t)nc * e;
  const double v[3] = {v0, v1, v2};
  double t = 0.0;
  bool any = false;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int o = a.dmi_ncomp > 0 ? a.dmi_off[3 * i + j] : 3 * i + j;
    if (o >= 0) {
      t = any ? fma(D[o], v[j], t) : __dmul_rn(D[o], v[j]);
      any = true;
    }
  }
  return t;
}

// Everything after the slot loop: the node's own terms, the outputs and the
// LLG RHS (shared by both forms of the field kernel).  The field kernels and
// their helpers compile with contraction off (__dmul_rn and __dadd_rn are a
// plain * and + in HIP, which the default fp-contract may fuse differently in
// the two kernels): only the written fma calls fuse, so both forms give the
// same bits.
__device__ __forceinline__ double mulr(double x, double y) { return __dmul_rn(x, y); }
__device__ __forceinline__ double addr(double x, double y) { return __dadd_rn(x, y); }
// a x b, each component as fma(a_i, b_j, -(a_j b_i))
__device__ __forceinline__ void cross3r(double ax, double ay, double az, double bx, double by,
                                        double bz, double& cx, double& cy, double& cz) {
  cx = fma(ay, bz, -mulr(az, by));
  cy = fma(az, bx, -mulr(ax, bz));
  cz = fma(ax, by, -mulr(ay, bx));
}

__device__ __forceinline__ void field_finish(const smx_field_args& a, int64_t k, bool stt, double hx,
                                             double hy, double hz, double gx, double gy, double gz,
                                             double zx, double zy, double zz) {
#pragma clang fp contract(off)
  const double mx = a.m[3 * k], my = a.m[3 * k + 1], mz = a.m[3 * k + 2];
  if (a.ani_pref != 0.0) {
    double nx, ny, nz;
    if (a.ani_axis) {
      nx = a.ani_axis[3 * k];
      ny = a.ani_axis[3 * k + 1];
      nz = a.ani_axis[3 * k + 2];
    } else {
      nx = a.ani_axis_u[0];
      ny = a.ani_axis_u[1];
      nz = a.ani_axis_u[2];
    }
    const double cdot = mulr(a.ani_pref, fma(mz, nz, fma(my, ny, mulr(mx, nx))));
    hx = fma(cdot, nx, hx);
    hy = fma(cdot, ny, hy);
    hz = fma(cdot, nz, hz);
  }
  if (a.cub_pref != 0.0) cubic_term(a, k, mx, my, mz, nullptr, hx, hy, hz);
  if (a.grad_vals) {
    const double iv = a.inv_vol[k];
    const double dx = mulr(gx, iv), dy = mulr(gy, iv), dz = mulr(gz, iv);
    if (a.h_demag) {
      a.h_demag[3 * k] = dx;
      a.h_demag[3 * k + 1] = dy;
      a.h_demag[3 * k + 2] = dz;
    }
    hx = addr(hx, dx);
    hy = addr(hy, dy);
    hz = addr(hz, dz);
  }
  if (a.h0) {
    hx = addr(hx, a.h0[3 * k]);
    hy = addr(hy, a.h0[3 * k + 1]);
    hz = addr(hz, a.h0[3 * k + 2]);
  } else {
    hx = addr(hx, a.h0_u[0]);
    hy = addr(hy, a.h0_u[1]);
    hz = addr(hz, a.h0_u[2]);
  }
  if (a.h_eff) {
    a.h_eff[3 * k] = hx;
    a.h_eff[3 * k + 1] = hy;
    a.h_eff[3 * k + 2] = hz;
  }
  if (a.rhs) {
    double mcx, mcy, mcz, mcmx, mcmy, mcmz;
    cross3r(mx, my, mz, hx, hy, hz, mcx, mcy, mcz);
    cross3r(mx, my, mz, mcx, mcy, mcz, mcmx, mcmy, mcmz);
    const double msq = fma(mz, mz, fma(my, my, mulr(mx, mx)));
    const double stab = mulr(a.stab, __dsub_rn(1.0, msq));
    double rx = fma(stab, mx, mulr(a.P, fma(a.alpha, mcmx, mulr(a.do_precess, mcx))));
    double ry = fma(stab, my, mulr(a.P, fma(a.alpha, mcmy, mulr(a.do_precess, mcy))));
    double rz = fma(stab, mz, mulr(a.P, fma(a.alpha, mcmz, mulr(a.do_precess, mcz))));
    if (stt) {  // - pref (a m x Z + b m x (m x Z))
      double c1x, c1y, c1z, c2x, c2y, c2z;
      cross3(mx, my, mz, zx, zy, zz, c1x, c1y, c1z);
      cross3(mx, my, mz, c1x, c1y, c1z, c2x, c2y, c2z);
      rx -= a.stt_pref * (a.stt_a * c1x + a.stt_b * c2x);
      ry -= a.stt_pref * (a.stt_a * c1y + a.stt_b * c2y);
      rz -= a.stt_pref * (a.stt_a * c1z + a.stt_b * c2z);
    }
    if (a.sot_fl != 0.0 || a.sot_dl != 0.0) {  // c_fl m x p + c_dl m x (m x p)
      double px, py, pz, c1x, c1y, c1z, c2x, c2y, c2z;
      sot_p(a, k, px, py, pz);
      cross3(mx, my, mz, px, py, pz, c1x, c1y, c1z);
      cross3(mx, my, mz, c1x, c1y, c1z, c2x, c2y, c2z);
      rx += a.sot_fl * c1x + a.sot_dl * c2x;
      ry += a.sot_fl * c1y + a.sot_dl * c2y;
      rz += a.sot_fl * c1z + a.sot_dl * c2z;
    }
    a.rhs[3 * k] = rx;
    a.rhs[3 * k + 1] = ry;
    a.rhs[3 * k + 2] = rz;
  }
}

__global__ __launch_bounds__(kBlock) void field_rhs_kernel(smx_field_args a) {
#pragma clang fp contract(off)
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.n) return;
  const int64_t s = k >> 6;
  const int lane = (int)(k & 63);
  const int64_t start = a.slice_ptr[s];
  const int64_t w = (a.slice_ptr[s + 1] - start) >> 6;
  const double shift = a.u1_sum ? (*a.u1_sum) / (double)a.n : 0.0;
  const bool stt = a.stt_vals && a.stt_pref != 0.0 && a.rhs;
  const bool need_m = a.ex_vals || a.dmi_vals || stt;
  double hx = 0.0, hy = 0.0, hz = 0.0;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  double zx = 0.0, zy = 0.0, zz = 0.0;
  for (int64_t j = 0; j < w; ++j) {
    const int64_t e = start + j * 64 + lane;
    const int64_t c = a.cols[e];
    if (need_m) {
      const double m0 = a.m[3 * c], m1 = a.m[3 * c + 1], m2 = a.m[3 * c + 2];
      if (a.ex_vals) {
        const double sv = a.ex_vals[e];
        hx = fma(sv, m0, hx);
        hy = fma(sv, m1, hy);
        hz = fma(sv, m2, hz);
      }
      if (stt) {
        const double tv = a.stt_vals[e];
        zx = fma(tv, m0, zx);
        zy = fma(tv, m1, zy);
        zz = fma(tv, m2, zz);
      }
      if (a.dmi_vals) {
        hx += dmi_row(a, e, 0, m0, m1, m2);
        hy += dmi_row(a, e, 1, m0, m1, m2);
        hz += dmi_row(a, e, 2, m0, m1, m2);
      }
    }
    if (a.grad_vals) {
      const double phi = (a.u1[c] - shift) + a.u2[c];
      gx = fma(a.grad_vals[3 * e], phi, gx);
      gy = fma(a.grad_vals[3 * e + 1], phi, gy);
      gz = fma(a.grad_vals[3 * e + 2], phi, gz);
    }
  }
  field_finish(a, k, stt, hx, hy, hz, gx, gy, gz, zx, zy, zz);
}

// The field kernel for the common term sets (exchange, DMI in one of the
// compact layouts below or none, demag gradient or none, uniaxial
// anisotropy; no cubic anisotropy or spin torques), with the term set and the DMI layout as template parameters.  The
// general kernel tests the optional terms inside its slot loop, so the
// compiler waits for each slot's loads at every test; here every load of U
// consecutive slots (columns, operator values, then the m and phi gathers) is
// issued before the first of their fmas, which run in the general kernel's
// order, so both kernels give the same bits.  DM: the stored DMI components
// as a mask of q = 3 i + j, stored in increasing q (0: no DMI).
template <int DM>
__device__ __forceinline__ double dmi_row_m(const double* D, int i, double v0, double v1, double v2) {
#pragma clang fp contract(off)
  const double v[3] = {v0, v1, v2};
  double t = 0.0;
  bool any = false;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int q = 3 * i + j;
    if ((DM >> q) & 1) {
      const int o = __builtin_popcount(DM & ((1 << q) - 1));
      t = any ? fma(D[o], v[j], t) : __dmul_rn(D[o], v[j]);
      any = true;
    }
  }
  return t;
}

// 202 005 nodes fill about three wavefronts per SIMD, so the kernel may use
// the registers of three (the loads of U slots in flight are its speed).
template <int DM, bool GRAD, int U>
__global__ __launch_bounds__(kBlock, 3) void field_rhs_fast_kernel(smx_field_args a) {
#pragma clang fp contract(off)
  constexpr int NC = __builtin_popcount(DM);
  constexpr int NCS = NC > 0 ? NC : 1;
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.n) return;
  const int64_t s = k >> 6;  // one slice per wavefront: w is wave-uniform
  const int lane = (int)(k & 63);
  const int64_t start = a.slice_ptr[s];
  const int64_t w = (a.slice_ptr[s + 1] - start) >> 6;
  const double shift = (GRAD && a.u1_sum) ? (*a.u1_sum) / (double)a.n : 0.0;
  double hx = 0.0, hy = 0.0, hz = 0.0;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (int64_t j0 = 0; j0 < w; j0 += U) {
    int32_t c[U];
    int64_t es[U];
    double ev[U], dv[U][NCS], gv[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // slots past the slice width re-read the last one; their sums are skipped
      es[u] = start + (j0 + u < w ? j0 + u : w - 1) * 64 + lane;
      c[u] = a.cols[es[u]];
    }
    // the columns first, then the operator values, then (once the columns
    // are in) the gathers: every wait below is for loads issued long before
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t e = es[u];
      ev[u] = a.ex_vals[e];
#pragma unroll
      for (int q = 0; q < NC; ++q) dv[u][q] = a.dmi_vals[(int64_t)NC * e + q];
      if (GRAD) {
#pragma unroll
        for (int q = 0; q < 3; ++q) gv[u][q] = a.grad_vals[3 * e + q];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    double m0[U], m1[U], m2[U], ph[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t cc = c[u];
      m0[u] = a.m[3 * cc];
      m1[u] = a.m[3 * cc + 1];
      m2[u] = a.m[3 * cc + 2];
      if (GRAD) ph[u] = (a.u1[cc] - shift) + a.u2[cc];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // a slot past the width keeps the sums by a select, not a branch: a
      // branch would let the compiler sink that slot's loads behind it
      const bool ok = u == 0 || j0 + u < w;
      double tx = fma(ev[u], m0[u], hx);
      double ty = fma(ev[u], m1[u], hy);
      double tz = fma(ev[u], m2[u], hz);
      if (NC > 0) {
        tx += dmi_row_m<DM>(dv[u], 0, m0[u], m1[u], m2[u]);
        ty += dmi_row_m<DM>(dv[u], 1, m0[u], m1[u], m2[u]);
        tz += dmi_row_m<DM>(dv[u], 2, m0[u], m1[u], m2[u]);
      }
      hx = ok ? tx : hx;
      hy = ok ? ty : hy;
      hz = ok ? tz : hz;
      if (GRAD) {
        const double sx = fma(gv[u][0], ph[u], gx);
        const double sy = fma(gv[u][1], ph[u], gy);
        const double sz = fma(gv[u][2], ph[u], gz);
        gx = ok ? sx : gx;
        gy = ok ? sy : gy;
        gz = ok ? sz : gz;
      }
    }
  }
  field_finish(a, k, false, hx, hy, hz, gx, gy, gz, 0.0, 0.0, 0.0);
}

// The DMI layout as a component mask when it is the compact layout (present
// components in increasing q) or the full one; -1 otherwise.
int dmi_mask(const smx_field_args& a) {
  if (!a.dmi_vals) return 0;
  if (a.dmi_ncomp == 0) return 0x1FF;
  int mask = 0, pos = 0;
  for (int q = 0; q < 9; ++q) {
    if (a.dmi_off[q] < 0) continue;
    if (a.dmi_off[q] != pos) return -1;
    mask |= 1 << q;
    ++pos;
  }
  return pos == a.dmi_ncomp ? mask : -1;
}

constexpr int kFieldU = 4;

template <int DM>
bool launch_field_fast_dm(const smx_field_args& a, hipStream_t st) {
  const dim3 g(smx::grid_for(a.n, kBlock)), b(kBlock);
  if (a.grad_vals) {
    hipLaunchKernelGGL((field_rhs_fast_kernel<DM, true, kFieldU>), g, b, 0, st, a);
  } else {
    hipLaunchKernelGGL((field_rhs_fast_kernel<DM, false, kFieldU>), g, b, 0, st, a);
  }
  return true;
}

// The fast form for exchange + (DMI in a listed layout) + (demag) +
// (uniaxial) without cubic anisotropy or spin torques; false: the general
// kernel must run.
bool launch_field_fast(const smx_field_args& a, hipStream_t st) {
  const bool stt = a.stt_vals && a.stt_pref != 0.0 && a.rhs;
  if (stt || !a.ex_vals || a.cub_pref != 0.0 || a.sot_fl != 0.0 || a.sot_dl != 0.0) return false;
  switch (dmi_mask(a)) {
    case 0: return launch_field_fast_dm<0>(a, st);
    case 0xE4: return launch_field_fast_dm<0xE4>(a, st);   // interfacial, n0 along z
    case 0x4E: return launch_field_fast_dm<0x4E>(a, st);   // interfacial, n0 along x
    case 0xAA: return launch_field_fast_dm<0xAA>(a, st);   // interfacial, n0 along y
    case 0xEE: return launch_field_fast_dm<0xEE>(a, st);   // bulk (no diagonal)
    case 0x1FF: return launch_field_fast_dm<0x1FF>(a, st);  // full blocks
    default: return false;
  }
}

// Everything of the JVP after the slot loop (shared by both forms of the JVP
// kernel; contraction off as in field_finish, so both give the same bits).
__device__ __forceinline__ void jvp_finish(const smx_field_args& a, int64_t k, const double* __restrict__ hm,
                                           const double* __restrict__ v, double* __restrict__ out,
                                           double gamma, bool stt, double hvx, double hvy, double hvz,
                                           double zmx, double zmy, double zmz, double zvx, double zvy,
                                           double zvz) {
#pragma clang fp contract(off)
  const double vx = v[3 * k], vy = v[3 * k + 1], vz = v[3 * k + 2];
  if (a.ani_pref != 0.0) {
    double nx, ny, nz;
    if (a.ani_axis) {
      nx = a.ani_axis[3 * k];
      ny = a.ani_axis[3 * k + 1];
      nz = a.ani_axis[3 * k + 2];
    } else {
      nx = a.ani_axis_u[0];
      ny = a.ani_axis_u[1];
      nz = a.ani_axis_u[2];
    }
    const double cdot = a.ani_pref * (vx * nx + vy * ny + vz * nz);
    hvx += cdot * nx;
    hvy += cdot * ny;
    hvz += cdot * nz;
  }
  const double mx = a.m[3 * k], my = a.m[3 * k + 1], mz = a.m[3 * k + 2];
  if (a.cub_pref != 0.0) {
    const double vk[3] = {vx, vy, vz};
    cubic_term(a, k, mx, my, mz, vk, hvx, hvy, hvz);
  }
  const double hx = hm[3 * k], hy = hm[3 * k + 1], hz = hm[3 * k + 2];
  const double cvx = vy * hz - vz * hy, cvy = vz * hx - vx * hz, cvz = vx * hy - vy * hx;
  const double cmx = my * hvz - mz * hvy, cmy = mz * hvx - mx * hvz, cmz = mx * hvy - my * hvx;
  const double mdHm = mx * hx + my * hy + mz * hz;
  const double mdHv = mx * hvx + my * hvy + mz * hvz;
  const double vdHm = vx * hx + vy * hy + vz * hz;
  const double mdv = mx * vx + my * vy + mz * vz;
  const double mdmm = mx * mx + my * my + mz * mz;
  const double common = vdHm + mdHv;
  const double dampx = vx * mdHm - 2.0 * hx * mdv + mx * common - hvx * mdmm;
  const double dampy = vy * mdHm - 2.0 * hy * mdv + my * common - hvy * mdmm;
  const double dampz = vz * mdHm - 2.0 * hz * mdv + mz * common - hvz * mdmm;
  const double coef = -gamma / (1.0 + a.alpha * a.alpha);
  const double sc = a.stab * (1.0 - mdmm);
  double ox = coef * (a.do_precess * (cvx + cmx) + a.alpha * dampx) + sc * vx -
              2.0 * a.stab * mx * mdv;
  double oy = coef * (a.do_precess * (cvy + cmy) + a.alpha * dampy) + sc * vy -
              2.0 * a.stab * my * mdv;
  double oz = coef * (a.do_precess * (cvz + cmz) + a.alpha * dampz) + sc * vz -
              2.0 * a.stab * mz * mdv;
  if (stt) {
    // d[m x Z] = v x Z + m x Zv;
    // d[m x (m x Z)] = v (m.Z) + m (v.Z + m.Zv) - Zv |m|^2 - 2 Z (m.v)
    double p1x, p1y, p1z, p2x, p2y, p2z;
    cross3(vx, vy, vz, zmx, zmy, zmz, p1x, p1y, p1z);
    cross3(mx, my, mz, zvx, zvy, zvz, p2x, p2y, p2z);
    const double mZ = mx * zmx + my * zmy + mz * zmz;
    const double vZ = vx * zmx + vy * zmy + vz * zmz;
    const double mZv = mx * zvx + my * zvy + mz * zvz;
    const double t = vZ + mZv;
    const double dx = vx * mZ + mx * t - zvx * mdmm - 2.0 * zmx * mdv;
    const double dy = vy * mZ + my * t - zvy * mdmm - 2.0 * zmy * mdv;
    const double dz = vz * mZ + mz * t - zvz * mdmm - 2.0 * zmz * mdv;
    ox -= a.stt_pref * (a.stt_a * (p1x + p2x) + a.stt_b * dx);
    oy -= a.stt_pref * (a.stt_a * (p1y + p2y) + a.stt_b * dy);
    oz -= a.stt_pref * (a.stt_a * (p1z + p2z) + a.stt_b * dz);
  }
  if (a.sot_fl != 0.0 || a.sot_dl != 0.0) {
    // d[m x p] = v x p;  d[m x (m x p)] = v x (m x p) + m x (v x p)
    double px, py, pz, cx, cy, cz, dcx, dcy, dcz, ex, ey, ez, fx, fy, fz;
    sot_p(a, k, px, py, pz);
    cross3(mx, my, mz, px, py, pz, cx, cy, cz);
    cross3(vx, vy, vz, px, py, pz, dcx, dcy, dcz);
    cross3(vx, vy, vz, cx, cy, cz, ex, ey, ez);
    cross3(mx, my, mz, dcx, dcy, dcz, fx, fy, fz);
    ox += a.sot_fl * dcx + a.sot_dl * (ex + fx);
    oy += a.sot_fl * dcy + a.sot_dl * (ey + fy);
    oz += a.sot_fl * dcz + a.sot_dl * (ez + fz);
  }
  out[3 * k] = ox;
  out[3 * k + 1] = oy;
  out[3 * k + 2] = oz;
}

__global__ __launch_bounds__(kBlock) void llg_jvp_kernel(smx_field_args a,
                                                         const double* __restrict__ hm,
                                                         const double* __restrict__ v,
                                                         double* __restrict__ out,
                                                         double gamma) {
#pragma clang fp contract(off)
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.n) return;
  const int64_t s = k >> 6;
  const int lane = (int)(k & 63);
  const int64_t start = a.slice_ptr[s];
  const int64_t w = (a.slice_ptr[s + 1] - start) >> 6;
  double hvx = 0.0, hvy = 0.0, hvz = 0.0;
  const bool stt = a.stt_vals && a.stt_pref != 0.0;
  double zmx = 0.0, zmy = 0.0, zmz = 0.0, zvx = 0.0, zvy = 0.0, zvz = 0.0;
  if (a.ex_vals || a.dmi_vals || stt) {
    for (int64_t j = 0; j < w; ++j) {
      const int64_t e = start + j * 64 + lane;
      const int64_t c = a.cols[e];
      const double v0 = v[3 * c], v1 = v[3 * c + 1], v2 = v[3 * c + 2];
      if (a.ex_vals) {
        const double sv = a.ex_vals[e];
        hvx = fma(sv, v0, hvx);
        hvy = fma(sv, v1, hvy);
        hvz = fma(sv, v2, hvz);
      }
      if (stt) {
        const double tv = a.stt_vals[e];
        zvx = fma(tv, v0, zvx);
        zvy = fma(tv, v1, zvy);
        zvz = fma(tv, v2, zvz);
        zmx = fma(tv, a.m[3 * c], zmx);
        zmy = fma(tv, a.m[3 * c + 1], zmy);
        zmz = fma(tv, a.m[3 * c + 2], zmz);
      }
      if (a.dmi_vals) {
        hvx += dmi_row(a, e, 0, v0, v1, v2);
        hvy += dmi_row(a, e, 1, v0, v1, v2);
        hvz += dmi_row(a, e, 2, v0, v1, v2);
      }
    }
  }
  jvp_finish(a, k, hm, v, out, gamma, stt, hvx, hvy, hvz, zmx, zmy, zmz, zvx, zvy, zvz);
}

// The JVP for exchange + DMI in a listed layout (or none) + uniaxial, without
// cubic anisotropy or spin torques: the slot loop of field_rhs_fast_kernel
// (U slots' column and value loads, then their v gathers, then the fmas in
// the general kernel's order), the same tail.
template <int DM, int U>
__global__ __launch_bounds__(kBlock, 3) void llg_jvp_fast_kernel(smx_field_args a,
                                                                 const double* __restrict__ hm,
                                                                 const double* __restrict__ v,
                                                                 double* __restrict__ out, double gamma) {
#pragma clang fp contract(off)
  constexpr int NC = __builtin_popcount(DM);
  constexpr int NCS = NC > 0 ? NC : 1;
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.n) return;
  const int64_t s = k >> 6;
  const int lane = (int)(k & 63);
  const int64_t start = a.slice_ptr[s];
  const int64_t w = (a.slice_ptr[s + 1] - start) >> 6;
  double hvx = 0.0, hvy = 0.0, hvz = 0.0;
  for (int64_t j0 = 0; j0 < w; j0 += U) {
    int32_t c[U];
    int64_t es[U];
    double ev[U], dv[U][NCS];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      es[u] = start + (j0 + u < w ? j0 + u : w - 1) * 64 + lane;
      c[u] = a.cols[es[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ev[u] = a.ex_vals[es[u]];
#pragma unroll
      for (int q = 0; q < NC; ++q) dv[u][q] = a.dmi_vals[(int64_t)NC * es[u] + q];
    }
    __builtin_amdgcn_sched_barrier(0);
    double v0[U], v1[U], v2[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t cc = c[u];
      v0[u] = v[3 * cc];
      v1[u] = v[3 * cc + 1];
      v2[u] = v[3 * cc + 2];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = u == 0 || j0 + u < w;
      double tx = fma(ev[u], v0[u], hvx);
      double ty = fma(ev[u], v1[u], hvy);
      double tz = fma(ev[u], v2[u], hvz);
      if (NC > 0) {
        tx += dmi_row_m<DM>(dv[u], 0, v0[u], v1[u], v2[u]);
        ty += dmi_row_m<DM>(dv[u], 1, v0[u], v1[u], v2[u]);
        tz += dmi_row_m<DM>(dv[u], 2, v0[u], v1[u], v2[u]);
      }
      hvx = ok ? tx : hvx;
      hvy = ok ? ty : hvy;
      hvz = ok ? tz : hvz;
    }
  }
  jvp_finish(a, k, hm, v, out, gamma, false, hvx, hvy, hvz, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
}

template <int DM>
bool launch_jvp_fast_dm(const smx_field_args& a, const double* hm, const double* v, double* out,
                        double gamma, hipStream_t st) {
  hipLaunchKernelGGL((llg_jvp_fast_kernel<DM, kFieldU>), dim3(smx::grid_for(a.n, kBlock)), dim3(kBlock), 0, st,
                     a, hm, v, out, gamma);
  return true;
}

// The fast JVP for the term sets of launch_field_fast; false: the general kernel.
bool launch_jvp_fast(const smx_field_args& a, const double* hm, const double* v, double* out, double gamma,
                     hipStream_t st) {
  const bool stt = a.stt_vals && a.stt_pref != 0.0;
  if (stt || !a.ex_vals || a.cub_pref != 0.0 || a.sot_fl != 0.0 || a.sot_dl != 0.0) return false;
  switch (dmi_mask(a)) {
    case 0: return launch_jvp_fast_dm<0>(a, hm, v, out, gamma, st);
    case 0xE4: return launch_jvp_fast_dm<0xE4>(a, hm, v, out, gamma, st);
    case 0x4E: return launch_jvp_fast_dm<0x4E>(a, hm, v, out, gamma, st);
    case 0xAA: return launch_jvp_fast_dm<0xAA>(a, hm, v, out, gamma, st);
    case 0xEE: return launch_jvp_fast_dm<0xEE>(a, hm, v, out, gamma, st);
    case 0x1FF: return launch_jvp_fast_dm<0x1FF>(a, hm, v, out, gamma, st);
    default: return false;
  }
}

__global__ __launch_bounds__(kBlock) void projdir_kernel(int64_t n, const double* __restrict__ m,
                                                         const double* __restrict__ h,
                                                         double* __restrict__ d) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const double mx = m[3 * k], my = m[3 * k + 1], mz = m[3 * k + 2];
  const double hx = h[3 * k], hy = h[3 * k + 1], hz = h[3 * k + 2];
  const double mh = mx * hx + my * hy + mz * hz;
  d[3 * k] = hx - mh * mx;
  d[3 * k + 1] = hy - mh * my;
  d[3 * k + 2] = hz - mh * mz;
}

// y = (I - C)^-1 (I + C) m, C v = c x v, c = (step/2) m x d:
// b = m + c x m, y = (b + c x b + c (c.b)) / (1 + |c|^2).
__global__ __launch_bounds__(kBlock) void cayley_kernel(int64_t n, const double* __restrict__ m,
                                                        const double* __restrict__ d, double step,
                                                        double* __restrict__ y) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const double mx = m[3 * k], my = m[3 * k + 1], mz = m[3 * k + 2];
  double cx, cy, cz;
  cross3(mx, my, mz, d[3 * k], d[3 * k + 1], d[3 * k + 2], cx, cy, cz);
  cx *= 0.5 * step;
  cy *= 0.5 * step;
  cz *= 0.5 * step;
  double tx, ty, tz;
  cross3(cx, cy, cz, mx, my, mz, tx, ty, tz);
  const double bx = mx + tx, by = my + ty, bz = mz + tz;
  double ux, uy, uz;
  cross3(cx, cy, cz, bx, by, bz, ux, uy, uz);
  const double cb = cx * bx + cy * by + cz * bz;
  const double inv = 1.0 / (1.0 + (cx * cx + cy * cy + cz * cz));
  y[3 * k] = (bx + ux + cx * cb) * inv;
  y[3 * k + 1] = (by + uy + cy * cb) * inv;
  y[3 * k + 2] = (bz + uz + cz * cb) * inv;
}

__global__ __launch_bounds__(kBlock) void normalize3_kernel(int64_t n, double* __restrict__ m) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const double x = m[3 * k], y = m[3 * k + 1], z = m[3 * k + 2];
  const double r = sqrt(x * x + y * y + z * z);
  const double s = r > 1e-300 ? 1.0 / r : 1.0;
  m[3 * k] = x * s;
  m[3 * k + 1] = y * s;
  m[3 * k + 2] = z * s;
}

// Per-block maxima of |v_i| into part[block]; the second kernel takes the max
// over the blocks (fixed, order-independent result).
__global__ __launch_bounds__(kBlock) void max_norm3_kernel(int64_t n, const double* __restrict__ v,
                                                           double* __restrict__ part) {
  __shared__ double sm[kBlock / 64];
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double r = 0.0;
  if (k < n) r = sqrt(v[3 * k] * v[3 * k] + v[3 * k + 1] * v[3 * k + 1] + v[3 * k + 2] * v[3 * k + 2]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sm[0];
    for (int i = 1; i < kBlock / 64; ++i) b = fmax(b, sm[i]);
    part[blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(kBlock) void max_reduce_kernel(int64_t g, const double* __restrict__ part,
                                                            double* __restrict__ out) {
  __shared__ double sm[kBlock / 64];
  double r = 0.0;
  for (int64_t i = threadIdx.x; i < g; i += kBlock) r = fmax(r, part[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sm[0];
    for (int i = 1; i < kBlock / 64; ++i) b = fmax(b, sm[i]);
    *out = b;
  }
}

__global__ __launch_bounds__(kBlock) void dot3_kernel(int64_t n, const double* __restrict__ a,
                                                      const double* __restrict__ b,
                                                      const double* __restrict__ w,
                                                      double* __restrict__ out,
                                                      double* __restrict__ red) {
  __shared__ double scratch[kBlock / 64];
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (k < n) {
    const double d = a[3 * k] * b[3 * k] + a[3 * k + 1] * b[3 * k + 1] + a[3 * k + 2] * b[3 * k + 2];
    v[0] = w ? w[k] * d : d;
  }
  if (smx::grid_sum_last<1>(v, red, scratch) && threadIdx.x == 0) *out = v[0];
}

struct RKArgs {
  const double* k[7];
  double c[7];
  double e[7];
  int ns;
};

__global__ __launch_bounds__(kBlock) void rk_combine_kernel(int64_t len, const double* __restrict__ y,
                                                            double dt, RKArgs A, double atol,
                                                            double rtol, double* __restrict__ out,
                                                            double* __restrict__ err_out,
                                                            double* __restrict__ red) {
  __shared__ double scratch[kBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (i < len) {
    double acc = 0.0, eacc = 0.0;
    for (int s = 0; s < A.ns; ++s) {
      const double ks = A.k[s][i];
      if (A.c[s] != 0.0) acc = fma(A.c[s], ks, acc);
      if (err_out && A.e[s] != 0.0) eacc = fma(A.e[s], ks, eacc);
    }
    const double yi = y[i];
    const double o = yi + dt * acc;
    out[i] = o;
    if (err_out) {
      const double ay = fabs(yi), ao = fabs(o);
      const double sc = atol + rtol * (ay > ao ? ay : ao);
      const double q = dt * eacc / sc;
      v[0] = q * q;
    }
  }
  if (err_out) {
    if (smx::grid_sum_last<1>(v, red, scratch) && threadIdx.x == 0) *err_out = sqrt(v[0] / (double)len);
  }
}

int check_field_args(const smx_field_args* a) {
  SMX_REQUIRE(a, SMX_ERR_VALUE, "null args");
  SMX_REQUIRE(a->n >= 0, SMX_ERR_SHAPE, "negative n");
  SMX_REQUIRE(a->m, SMX_ERR_VALUE, "m is required");
  const bool graph = a->ex_vals || a->dmi_vals || a->grad_vals;
  SMX_REQUIRE(!graph || (a->slice_ptr && a->cols), SMX_ERR_VALUE, "graph pattern missing");
  SMX_REQUIRE(!a->grad_vals || (a->u1 && a->u2 && a->inv_vol), SMX_ERR_VALUE,
              "demag gradient needs u1, u2 and inv_vol");
  return SMX_OK;
}

}  // namespace

extern "C" int smx_version(void) { return 100; }
extern "C" const char* smx_last_error(void) { return smx::g_err; }

extern "C" int smx_field_llg_rhs(const smx_field_args* a, void* stream) {
  int rc = check_field_args(a);
  if (rc) return rc;
  if (a->n == 0) return SMX_OK;
  if (!launch_field_fast(*a, smx::as_stream(stream))) {
    hipLaunchKernelGGL(field_rhs_kernel, dim3(smx::grid_for(a->n, kBlock)), dim3(kBlock), 0,
                       smx::as_stream(stream), *a);
  }
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_llg_jvp(const smx_field_args* a, const double* hm, const double* v,
                           double* out, double gamma, void* stream) {
  int rc = check_field_args(a);
  if (rc) return rc;
  SMX_REQUIRE(hm && v && out, SMX_ERR_VALUE, "null pointer argument");
  if (a->n == 0) return SMX_OK;
  if (!launch_jvp_fast(*a, hm, v, out, gamma, smx::as_stream(stream))) {
    hipLaunchKernelGGL(llg_jvp_kernel, dim3(smx::grid_for(a->n, kBlock)), dim3(kBlock), 0,
                       smx::as_stream(stream), *a, hm, v, out, gamma);
  }
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_projected_direction(int64_t n, const double* m, const double* h, double* d,
                                       void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative n");
  SMX_REQUIRE(m && h && d, SMX_ERR_VALUE, "null pointer argument");
  if (n == 0) return SMX_OK;
  hipLaunchKernelGGL(projdir_kernel, dim3(smx::grid_for(n, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), n, m, h, d);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_cayley_step(int64_t n, const double* m, const double* d, double step, double* out,
                               void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative n");
  SMX_REQUIRE(m && d && out, SMX_ERR_VALUE, "null pointer argument");
  if (n == 0) return SMX_OK;
  hipLaunchKernelGGL(cayley_kernel, dim3(smx::grid_for(n, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), n, m, d, step, out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_normalize3(int64_t n, double* m, void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative n");
  SMX_REQUIRE(m, SMX_ERR_VALUE, "null pointer argument");
  if (n == 0) return SMX_OK;
  hipLaunchKernelGGL(normalize3_kernel, dim3(smx::grid_for(n, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), n, m);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_max_norm3(int64_t n, const double* v, double* out, double* red, void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative n");
  SMX_REQUIRE(v && out && red, SMX_ERR_VALUE, "null pointer argument");
  hipStream_t st = smx::as_stream(stream);
  const int g = smx::grid_for(n, kBlock);
  if (n > 0) {
    hipLaunchKernelGGL(max_norm3_kernel, dim3(g), dim3(kBlock), 0, st, n, v, red + 8);
    SMX_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(kBlock), 0, st, n > 0 ? (int64_t)g : 0,
                     red + 8, out);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_dot3_weighted(int64_t n, const double* a, const double* b, const double* w,
                                 double* out, double* red, void* stream) {
  SMX_REQUIRE(n >= 0, SMX_ERR_SHAPE, "negative n");
  SMX_REQUIRE(a && b && out && red, SMX_ERR_VALUE, "null pointer argument");
  hipLaunchKernelGGL(dot3_kernel, dim3(smx::grid_for(n, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), n, a, b, w, out, red);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_rk_combine(int64_t len, const double* y, double dt, int32_t nstage,
                              const double* const* k, const double* c, const double* e,
                              double atol, double rtol, double* out, double* err_out,
                              double* red, void* stream) {
  SMX_REQUIRE(len >= 0, SMX_ERR_SHAPE, "negative length");
  SMX_REQUIRE(nstage >= 0 && nstage <= 7, SMX_ERR_VALUE, "nstage must be in [0, 7]");
  SMX_REQUIRE(y && out && (nstage == 0 || (k && c)), SMX_ERR_VALUE, "null pointer argument");
  SMX_REQUIRE(!err_out || (e && red), SMX_ERR_VALUE, "err_out needs e and red");
  RKArgs A;
  A.ns = nstage;
  for (int s = 0; s < 7; ++s) {
    A.k[s] = s < nstage ? k[s] : nullptr;
    A.c[s] = s < nstage ? c[s] : 0.0;
    A.e[s] = (s < nstage && e) ? e[s] : 0.0;
  }
  for (int s = 0; s < nstage; ++s) SMX_REQUIRE(A.k[s], SMX_ERR_VALUE, "null stage %d", s);
  if (len == 0) return SMX_OK;
  hipLaunchKernelGGL(rk_combine_kernel, dim3(smx::grid_for(len, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), len, y, dt, A, atol, rtol, out, err_out, red);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}
