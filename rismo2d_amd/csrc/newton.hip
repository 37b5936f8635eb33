// Newton update on the device (EQS_UVS2D::Execute, EqsUVS2D.cpp:398-760 and EQS::Update,
// Update.cpp:42-187): correction statistics, relaxation factors, correction applied in
// the rotated frame of rotated nodes, velocity cap, depth / marsh handling and the
// midside water level and bed by linear interpolation.
#include <cmath>

#include "rmx_internal.h"

namespace rmx {

constexpr int UB = 256;          // blocks of the statistics reduction

// per block: max|dU|,max|dV|,max|dh|, sum|dU|,sum|dV|,sum|dh|, cnt U,V,h, max|dUs|, max|dS|
__global__ void __launch_bounds__(256) k_update_stats(int np, const int* __restrict__ eqno,
                                                      const double* __restrict__ X,
                                                      double* __restrict__ part)
{
  __shared__ double sh[11][256];
  double v[11];
  for (int k = 0; k < 11; k++) v[k] = 0.0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < np; n += gridDim.x * blockDim.x) {
    double d[3] = { 0.0, 0.0, 0.0 };
    for (int q = 0; q < 3; q++) {
      const int e = eqno[3 * (size_t)n + q];
      if (e < 0) continue;
      const double a = fabs(X[e]);
      d[q] = a;
      v[q] = fmax(v[q], a);
      v[3 + q] += a;
      v[6 + q] += 1.0;
    }
    v[9] = fmax(v[9], sqrt(d[0] * d[0] + d[1] * d[1]));
    v[10] = fmax(v[10], d[2]);
  }
  for (int k = 0; k < 11; k++) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 11; k++) {
        const double a = sh[k][threadIdx.x], b = sh[k][threadIdx.x + s];
        sh[k][threadIdx.x] = (k < 3 || k >= 9) ? fmax(a, b) : a + b;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 11; k++) part[k * UB + blockIdx.x] = sh[k][0];
}

struct UpdParams {
  double relaxUs, relaxS, maxUs, dryLimit, hmin;
  int dryMethod;
};

__global__ void k_apply(int np, const int* __restrict__ eqno, const double* __restrict__ X,
                        const uint32_t* __restrict__ flag, const double* __restrict__ rotm,
                        const double* __restrict__ zor, double* __restrict__ U,
                        double* __restrict__ V, double* __restrict__ S, double* __restrict__ z,
                        UpdParams P)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  const int eU = eqno[3 * (size_t)n], eV = eqno[3 * (size_t)n + 1], eS = eqno[3 * (size_t)n + 2];
  const double dU = eU >= 0 ? X[eU] * P.relaxUs : 0.0;
  const double dV = eV >= 0 ? X[eV] * P.relaxUs : 0.0;
  const double dS = eS >= 0 ? X[eS] * P.relaxS : 0.0;
  double u = U[n], v = V[n];
  if (flag[n] & dRotat) {
    const double* R = rotm + 4 * (size_t)n;
    u += R[0] * dU; v += R[2] * dU;
    u += R[1] * dV; v += R[3] * dV;
  } else {
    u += dU; v += dV;
  }
  const double us = sqrt(u * u + v * v);
  if (us > P.maxUs) { u = u * P.maxUs / us; v = v * P.maxUs / us; }
  U[n] = u; V[n] = v;
  if (eS >= 0) {
    double s = S[n] + dS;
    if (P.dryMethod == 0 || P.dryMethod == 2 || P.dryMethod == 3) {
      z[n] = (s - zor[n] < P.dryLimit) ? s - P.dryLimit : zor[n];
    } else if (s - z[n] <= 0.0) {
      s = z[n] + P.hmin;
    }
    S[n] = s;
  }
}

__global__ void k_midside(int ne, const int* __restrict__ elem, const int* __restrict__ nnd,
                          double* __restrict__ S, double* __restrict__ z)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int k = nnd[e], ncn = k == 6 ? 3 : 4;
  const int* nd = elem + 8 * (size_t)e;
  for (int i = ncn; i < k; i++) {
    const int l = nd[i - ncn], r = nd[(i - ncn + 1) % ncn];
    S[nd[i]] = 0.5 * (S[l] + S[r]);
    z[nd[i]] = 0.5 * (z[l] + z[r]);
  }
}

// ---- theta scheme (EQS_UVS2D::Predict / Timegrad, EqsUVS2D.cpp:1101-1230) ----

// Predict: current time derivatives zero; where the previous level lies below
// zor + dryLimit the previous derivatives are zero and, at the first time step or where the
// node was dry at the previous time step, the previous state is the current one.
__global__ void k_predict(int np, int first, double dryLimit, const uint8_t* __restrict__ nwd,
                          const double* __restrict__ zor,
                          const double* __restrict__ U, const double* __restrict__ V,
                          const double* __restrict__ S, double* __restrict__ dUdt,
                          double* __restrict__ dVdt, double* __restrict__ dSdt,
                          double* __restrict__ Uo, double* __restrict__ Vo,
                          double* __restrict__ So, double* __restrict__ dUdto,
                          double* __restrict__ dVdto, double* __restrict__ dSdto)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  dUdt[n] = 0.0; dVdt[n] = 0.0; dSdt[n] = 0.0;
  if (So[n] < zor[n] + dryLimit) {
    if (first || (nwd[n] & WD_DRYPREV)) { Uo[n] = U[n]; Vo[n] = V[n]; So[n] = S[n]; }
    dUdto[n] = 0.0; dVdto[n] = 0.0; dSdto[n] = 0.0;
  }
}

// Timegrad: d/dt = (1 - 1/theta) d/dt_old + (v - v_old) / (dt theta); zero at nodes dry or
// marsh now or at the previous time step
__global__ void k_timegrad(int np, double iTheta, double thdt, const uint8_t* __restrict__ nwd,
                           const double* __restrict__ U,
                           const double* __restrict__ V, const double* __restrict__ S,
                           const double* __restrict__ Uo, const double* __restrict__ Vo,
                           const double* __restrict__ So, const double* __restrict__ dUdto,
                           const double* __restrict__ dVdto, const double* __restrict__ dSdto,
                           double* __restrict__ dUdt, double* __restrict__ dVdt,
                           double* __restrict__ dSdt)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  if (nwd[n] & (WD_DRY | WD_MARSH | WD_DRYPREV | WD_MARSHPREV)) {
    dUdt[n] = 0.0; dVdt[n] = 0.0; dSdt[n] = 0.0;
    return;
  }
  dUdt[n] = iTheta * dUdto[n] + thdt * (U[n] - Uo[n]);
  dVdt[n] = iTheta * dVdto[n] + thdt * (V[n] - Vo[n]);
  dSdt[n] = iTheta * dSdto[n] + thdt * (S[n] - So[n]);
}

// GRID::SetSlipFlow (SlipFlow.cpp:43-78): at rotated nodes the velocity and its time
// derivative keep only their component along the first column of the rotation (the
// tangent of a wall, the inflow or outflow normal of an open boundary)
__global__ void k_slip_flow(int np, const uint32_t* __restrict__ flag, const double* __restrict__ rotm,
                            double* __restrict__ U, double* __restrict__ V,
                            double* __restrict__ dUdt, double* __restrict__ dVdt)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np || !(flag[n] & dRotat)) return;
  const double r00 = rotm[4 * (size_t)n], r10 = rotm[4 * (size_t)n + 2];
  double t = r00 * U[n] + r10 * V[n];
  U[n] = r00 * t;
  V[n] = r10 * t;
  t = r00 * dUdt[n] + r10 * dVdt[n];
  dUdt[n] = r00 * t;
  dVdt[n] = r10 * t;
}

void dev_slip_flow(Ctx& c)
{
  const int np = c.hm.np;
  k_slip_flow<<<(np + 255) / 256, 256, 0, c.stream>>>(np, c.dflag.p, c.rotm.p, c.U.p, c.V.p,
                                                      c.dUdt.p, c.dVdt.p);
  HIPCHECK(hipGetLastError());
}

void dev_time_init(Ctx& c)
{
  const size_t b = sizeof(double) * (size_t)c.hm.np;
  HIPCHECK(hipMemcpyAsync(c.Uo.p, c.U.p, b, hipMemcpyDeviceToDevice, c.stream));
  HIPCHECK(hipMemcpyAsync(c.Vo.p, c.V.p, b, hipMemcpyDeviceToDevice, c.stream));
  HIPCHECK(hipMemcpyAsync(c.So.p, c.S.p, b, hipMemcpyDeviceToDevice, c.stream));
  for (double* p : {c.dUdt.p, c.dVdt.p, c.dSdt.p, c.dUdto.p, c.dVdto.p, c.dSdto.p})
    HIPCHECK(hipMemsetAsync(p, 0, b, c.stream));
}

void dev_time_step(Ctx& c, int first)
{
  const int np = c.hm.np;
  k_predict<<<(np + 255) / 256, 256, 0, c.stream>>>(np, first, c.par.dry_limit, c.d_nwd.p,
      c.zor.p, c.U.p, c.V.p, c.S.p, c.dUdt.p, c.dVdt.p, c.dSdt.p, c.Uo.p, c.Vo.p, c.So.p,
      c.dUdto.p, c.dVdto.p, c.dSdto.p);
  HIPCHECK(hipGetLastError());
}

void dev_time_commit(Ctx& c)
{
  const size_t b = sizeof(double) * (size_t)c.hm.np;
  HIPCHECK(hipMemcpyAsync(c.Uo.p, c.U.p, b, hipMemcpyDeviceToDevice, c.stream));
  HIPCHECK(hipMemcpyAsync(c.Vo.p, c.V.p, b, hipMemcpyDeviceToDevice, c.stream));
  HIPCHECK(hipMemcpyAsync(c.So.p, c.S.p, b, hipMemcpyDeviceToDevice, c.stream));
  HIPCHECK(hipMemcpyAsync(c.dUdto.p, c.dUdt.p, b, hipMemcpyDeviceToDevice, c.stream));
  HIPCHECK(hipMemcpyAsync(c.dVdto.p, c.dVdt.p, b, hipMemcpyDeviceToDevice, c.stream));
  HIPCHECK(hipMemcpyAsync(c.dSdto.p, c.dSdt.p, b, hipMemcpyDeviceToDevice, c.stream));
}

// Start of a flow cycle (EQS_UVS2D::Execute, EqsUVS2D.cpp:97-147): relaxation methods 3 and
// 4 start the relaxed time steps at relaxTimeFlow (at most the time step); the Jacobian's
// time terms take 1 / dt_r (stationary) or 1 / dt_r / theta, else 1 / dt / theta (unsteady)
// or none.
void cycle_start(Ctx& c)
{
  const double dt = c.unsteady ? c.dt : c.par.interval;
  const double rt = c.par.relax_time;
  c.dt_uv = c.dt_h = c.last_dt = rt > 0.0 && rt <= dt ? rt : dt;
  if (c.par.relax_method >= 3) {
    const double th = c.unsteady ? c.theta : 1.0;
    c.thdt_uv = 1.0 / c.dt_uv / th;
    c.thdt_h = 1.0 / c.dt_h / th;
  } else {
    c.thdt_uv = c.thdt_h = c.unsteady ? 1.0 / c.dt / c.theta : 0.0;
  }
}

void dev_newton_update(Ctx& c, rmx_step_stats& st)
{
  const int np = c.hm.np;
  double* part = c.red.p + 64;
  k_update_stats<<<UB, 256, 0, c.stream>>>(np, c.eqno.p, c.X.p, part);
  HIPCHECK(hipGetLastError());
  std::vector<double> h(11 * UB);
  HIPCHECK(hipMemcpyAsync(h.data(), part, sizeof(double) * 11 * UB, hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  double v[11];
  for (int k = 0; k < 11; k++) {
    double a = 0.0;
    for (int b = 0; b < UB; b++) {
      const double x = h[k * UB + b];
      a = (k < 3 || k >= 9) ? std::fmax(a, x) : a + x;
    }
    v[k] = a;
  }
  st.nrconv = 1;
  for (int q = 0; q < 3; q++) {
    st.maxAbs[q] = v[q];
    st.avAbs[q] = v[6 + q] > 0 ? v[3 + q] / v[6 + q] : 0.0;
    if (std::fabs(v[q]) > (q < 2 ? c.par.conv_uv : c.par.conv_s)) st.nrconv = 0;
  }
  double relaxUs = 1.0, relaxS = 1.0;
  st.relax_time_last = c.dt_h;
  if (c.par.relax_method >= 3) {
    // the relaxed time steps (EqsUVS2D.cpp:540-627): no convergence before they reach the
    // time step; each grows by the ratio of the maximum change to the largest correction,
    // within [relaxTimeFlow, dt], the two kept equal
    const double dt = c.unsteady ? c.dt : c.par.interval;
    const double rt = c.par.relax_time > 0.0 && c.par.relax_time <= dt ? c.par.relax_time : dt;
    if (c.dt_h < dt || c.dt_uv < dt) st.nrconv = 0;
    relaxUs = c.par.max_delta_uv / v[9];
    relaxS = c.par.max_delta_s / v[10];
    c.last_dt = c.dt_h;
    c.dt_uv *= relaxUs;
    c.dt_h *= relaxS;
    if (c.dt_uv > dt) c.dt_uv = dt;
    if (c.dt_uv < rt) c.dt_uv = rt;
    if (c.dt_h > dt) c.dt_h = dt;
    if (c.dt_h < rt) c.dt_h = rt;
    if (c.dt_h < c.dt_uv) c.dt_uv = c.dt_h; else c.dt_h = c.dt_uv;
    const double th = c.unsteady ? c.theta : 1.0;
    c.thdt_uv = 1.0 / c.dt_uv / th;
    c.thdt_h = 1.0 / c.dt_h / th;
    if (relaxUs > 1.0) relaxUs = 1.0;
    if (relaxS > 1.0) relaxS = 1.0;
  } else if (c.par.relax_method == 2) {
    relaxUs = c.par.max_delta_uv / v[9];
    relaxS = c.par.max_delta_s / v[10];
    if (relaxUs > 1.0) relaxUs = 1.0;
    if (relaxS > 1.0) relaxS = 1.0;
  }
  st.relaxUs = relaxUs; st.relaxS = relaxS;
  st.relax_time = c.dt_h;
  UpdParams P;
  P.relaxUs = relaxUs; P.relaxS = relaxS; P.maxUs = c.par.max_us; P.dryLimit = c.par.dry_limit;
  P.hmin = c.par.hmin(); P.dryMethod = c.par.dry_method;
  k_apply<<<(np + 255) / 256, 256, 0, c.stream>>>(np, c.eqno.p, c.X.p, c.dflag.p, c.rotm.p,
                                                  c.zor.p, c.U.p, c.V.p, c.S.p, c.z.p, P);
  k_midside<<<(c.hm.ne + 255) / 256, 256, 0, c.stream>>>(c.hm.ne, c.d_elem.p, c.d_nnd.p, c.S.p,
                                                         c.z.p);
  HIPCHECK(hipGetLastError());
  if (c.unsteady) {                       // time derivatives of the corrected state
    k_timegrad<<<(np + 255) / 256, 256, 0, c.stream>>>(np, 1.0 - 1.0 / c.theta,
        1.0 / c.dt / c.theta, c.d_nwd.p, c.U.p, c.V.p, c.S.p, c.Uo.p, c.Vo.p, c.So.p, c.dUdto.p,
        c.dVdto.p, c.dSdto.p, c.dUdt.p, c.dVdt.p, c.dSdt.p);
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipStreamSynchronize(c.stream));
}

}  // namespace rmx
