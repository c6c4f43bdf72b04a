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

__global__ __launch_bounds__(kBlock) void field_rhs_kernel(smx_field_args a) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.n) return;
  const int64_t s = k >> 6;
  const int lane = (int)(k & 63);
  const int64_t start = a.slice_ptr[s];
  const int64_t w = (a.slice_ptr[s + 1] - start) >> 6;
  const double shift = a.u1_sum ? (*a.u1_sum) / (double)a.n : 0.0;
  const bool need_m = a.ex_vals || a.dmi_vals;
  double hx = 0.0, hy = 0.0, hz = 0.0;
  double gx = 0.0, gy = 0.0, gz = 0.0;
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
      if (a.dmi_vals) {
        const double* D = a.dmi_vals + 9 * e;
        hx += D[0] * m0 + D[1] * m1 + D[2] * m2;
        hy += D[3] * m0 + D[4] * m1 + D[5] * m2;
        hz += D[6] * m0 + D[7] * m1 + D[8] * m2;
      }
    }
    if (a.grad_vals) {
      const double phi = (a.u1[c] - shift) + a.u2[c];
      gx = fma(a.grad_vals[3 * e], phi, gx);
      gy = fma(a.grad_vals[3 * e + 1], phi, gy);
      gz = fma(a.grad_vals[3 * e + 2], phi, gz);
    }
  }
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
    const double cdot = a.ani_pref * (mx * nx + my * ny + mz * nz);
    hx += cdot * nx;
    hy += cdot * ny;
    hz += cdot * nz;
  }
  if (a.grad_vals) {
    const double iv = a.inv_vol[k];
    const double dx = gx * iv, dy = gy * iv, dz = gz * iv;
    if (a.h_demag) {
      a.h_demag[3 * k] = dx;
      a.h_demag[3 * k + 1] = dy;
      a.h_demag[3 * k + 2] = dz;
    }
    hx += dx;
    hy += dy;
    hz += dz;
  }
  if (a.h0) {
    hx += a.h0[3 * k];
    hy += a.h0[3 * k + 1];
    hz += a.h0[3 * k + 2];
  } else {
    hx += a.h0_u[0];
    hy += a.h0_u[1];
    hz += a.h0_u[2];
  }
  if (a.h_eff) {
    a.h_eff[3 * k] = hx;
    a.h_eff[3 * k + 1] = hy;
    a.h_eff[3 * k + 2] = hz;
  }
  if (a.rhs) {
    const double mcx = my * hz - mz * hy;
    const double mcy = mz * hx - mx * hz;
    const double mcz = mx * hy - my * hx;
    const double mcmx = my * mcz - mz * mcy;
    const double mcmy = mz * mcx - mx * mcz;
    const double mcmz = mx * mcy - my * mcx;
    const double stab = a.stab * (1.0 - (mx * mx + my * my + mz * mz));
    a.rhs[3 * k] = a.P * (a.do_precess * mcx + a.alpha * mcmx) + stab * mx;
    a.rhs[3 * k + 1] = a.P * (a.do_precess * mcy + a.alpha * mcmy) + stab * my;
    a.rhs[3 * k + 2] = a.P * (a.do_precess * mcz + a.alpha * mcmz) + stab * mz;
  }
}

__global__ __launch_bounds__(kBlock) void llg_jvp_kernel(smx_field_args a,
                                                         const double* __restrict__ hm,
                                                         const double* __restrict__ v,
                                                         double* __restrict__ out,
                                                         double gamma) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.n) return;
  const int64_t s = k >> 6;
  const int lane = (int)(k & 63);
  const int64_t start = a.slice_ptr[s];
  const int64_t w = (a.slice_ptr[s + 1] - start) >> 6;
  double hvx = 0.0, hvy = 0.0, hvz = 0.0;
  if (a.ex_vals || a.dmi_vals) {
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
      if (a.dmi_vals) {
        const double* D = a.dmi_vals + 9 * e;
        hvx += D[0] * v0 + D[1] * v1 + D[2] * v2;
        hvy += D[3] * v0 + D[4] * v1 + D[5] * v2;
        hvz += D[6] * v0 + D[7] * v1 + D[8] * v2;
      }
    }
  }
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
  out[3 * k] = coef * (a.do_precess * (cvx + cmx) + a.alpha * dampx) + sc * vx -
               2.0 * a.stab * mx * mdv;
  out[3 * k + 1] = coef * (a.do_precess * (cvy + cmy) + a.alpha * dampy) + sc * vy -
                   2.0 * a.stab * my * mdv;
  out[3 * k + 2] = coef * (a.do_precess * (cvz + cmz) + a.alpha * dampz) + sc * vz -
                   2.0 * a.stab * mz * mdv;
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
  hipLaunchKernelGGL(field_rhs_kernel, dim3(smx::grid_for(a->n, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), *a);
  SMX_LAUNCH_CHECK();
  return SMX_OK;
}

extern "C" int smx_llg_jvp(const smx_field_args* a, const double* hm, const double* v,
                           double* out, double gamma, void* stream) {
  int rc = check_field_args(a);
  if (rc) return rc;
  SMX_REQUIRE(hm && v && out, SMX_ERR_VALUE, "null pointer argument");
  if (a->n == 0) return SMX_OK;
  hipLaunchKernelGGL(llg_jvp_kernel, dim3(smx::grid_for(a->n, kBlock)), dim3(kBlock), 0,
                     smx::as_stream(stream), *a, hm, v, out, gamma);
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
