// Krylov loops on the device: right-preconditioned BiCGStab (P_BCGSTABD::Iterate,
// P_bcgstabd.cpp:98-426) and flexible GMRES(m) with classical Gram-Schmidt and Givens
// rotations (P_FGMRESD::Iterate, P_fgmresd.cpp:90-389, with the Hessenberg entries
// assigned from the Gram-Schmidt dot products).  Vectors are FP64 and stay on the
// device; inner products are FP64 with a fixed-order two-stage reduction; only the
// handful of scalars per iteration (and FGMRES's small Hessenberg system) cross to the
// host.  Around either loop an optional FP64 iterative refinement solves with the
// FP32 matrix and factors and corrects with the FP64 residual.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "rmx_internal.h"

namespace rmx {


enum { S_RHO = 0, S_RV = 2, S_TS = 3, S_TT = 4, S_RR = 5, S_B2 = 6 };

__global__ void k_copy3(long long n, const double* __restrict__ b, double* __restrict__ r,
                        double* __restrict__ r0, double* __restrict__ p, double* __restrict__ x)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double v = b[i];
    r[i] = v; r0[i] = v; p[i] = v; x[i] = 0.0;
  }
}

__global__ void k_bicg_s(long long n, const double* __restrict__ sc, int rho, const double* __restrict__ r,
                         const double* __restrict__ v, double* __restrict__ s)
{
  const double alpha = sc[rho] / sc[S_RV];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s[i] = r[i] + (-alpha) * v[i];
}

__global__ void k_bicg_xr(long long n, const double* __restrict__ sc, int rho,
                          const double* __restrict__ p, const double* __restrict__ s,
                          const double* __restrict__ t, double* __restrict__ x,
                          double* __restrict__ r)
{
  const double alpha = sc[rho] / sc[S_RV];
  const double omega = sc[S_TS] / sc[S_TT];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double xi = x[i] + alpha * p[i];
    xi += omega * s[i];
    x[i] = xi;
    r[i] = s[i] + (-omega) * t[i];
  }
}

__global__ void k_bicg_p(long long n, const double* __restrict__ sc, int rho_old, int rho_new,
                         const double* __restrict__ r, const double* __restrict__ v,
                         double* __restrict__ p)
{
  const double alpha = sc[rho_old] / sc[S_RV];
  const double omega = sc[S_TS] / sc[S_TT];
  const double beta = sc[rho_new] * alpha / (sc[rho_old] * omega);
  const double bo = -beta * omega;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double pi = beta * p[i];
    pi += r[i];
    pi += bo * v[i];
    p[i] = pi;
  }
}

__global__ void k_copy(long long n, const double* __restrict__ a, double* __restrict__ b)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void k_axpy(long long n, double a, const double* __restrict__ x, double* __restrict__ y)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

// x *= 1 / sqrt(sc[slot]) (scale from the device), or x *= a when sc is null; with x32 the
// scaled vector is also stored in FP32 (the Gram-Schmidt copy of the basis)
__global__ void k_scal_dev(long long n, const double* __restrict__ sc, int slot, double a,
                           double* __restrict__ x, float* __restrict__ x32)
{
  const double t = sc ? 1.0 / sqrt(sc[slot]) : a;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double v = x[i] * t;
    x[i] = v;
    if (x32) x32[i] = (float)v;
  }
}

// y -= sum_j h[j] * V[j]   (classical Gram-Schmidt update, h on the device)
__global__ void k_gs_update(long long n, int K, const double* __restrict__ h,
                            const double* __restrict__ Vb, long long ld, double* __restrict__ y)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double v = y[i];
    for (int j = 0; j < K; j++) v += (-h[j]) * Vb[j * ld + i];
    y[i] = v;
  }
}

// x += sum_j rs[j] * Z[j]
__global__ void k_comb(long long n, int K, const double* __restrict__ rs,
                       const double* __restrict__ Zb, long long ld, double* __restrict__ x)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double v = x[i];
    for (int j = 0; j < K; j++) v += rs[j] * Zb[j * ld + i];
    x[i] = v;
  }
}

constexpr int RED_BLOCKS = 512;

__device__ __forceinline__ double bsum(double v, double* sh)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += sh[i];
  __syncthreads();
  return s;
}

// partial[j*MDOT_BLOCKS + block] = <V[j], y> for j < K <= 32: y read once, K accumulators
// per thread; enough workgroups for eight waves per SIMD
constexpr int MDOT_MAX = 32;
constexpr int MDOT_BLOCKS = 2048;
__global__ void __launch_bounds__(256) k_mdot(long long n, int K, const double* __restrict__ Vb,
                                              long long ld, const double* __restrict__ y,
                                              double* __restrict__ partial)
{
  __shared__ double sh[4];
  double acc[MDOT_MAX];
#pragma unroll
  for (int j = 0; j < MDOT_MAX; j++) acc[j] = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double yi = y[i];
#pragma unroll
    for (int j = 0; j < MDOT_MAX; j++)
      if (j < K) acc[j] += Vb[j * ld + i] * yi;
  }
#pragma unroll
  for (int j = 0; j < MDOT_MAX; j++) {
    if (j < K) {
      const double s = bsum(acc[j], sh);
      if (threadIdx.x == 0) partial[j * MDOT_BLOCKS + blockIdx.x] = s;
    }
  }
}

// FP32 basis (leading dimension ld a multiple of 4, zero padding past n): four entries per
// thread with 16-byte loads.  k_mdot32: partial[j*MDOT_BLOCKS + block] = <V[j], y>;
// k_gs_update32: y -= sum_j h[j] V[j] and partial[block] = the block's part of <y, y> (the
// norm of the updated vector in the same pass).  Both run on exactly MDOT_BLOCKS workgroups.
__global__ void __launch_bounds__(256) k_mdot32(long long n, int K, const float* __restrict__ Vb,
                                                long long ld, const double* __restrict__ y,
                                                double* __restrict__ partial)
{
  __shared__ double sh[4];
  double acc[MDOT_MAX];
#pragma unroll
  for (int j = 0; j < MDOT_MAX; j++) acc[j] = 0.0;
  const long long nq = (n + 3) / 4;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += (long long)gridDim.x * blockDim.x) {
    const long long i = 4 * q;
    const double y0 = y[i];
    const double y1 = i + 1 < n ? y[i + 1] : 0.0;
    const double y2 = i + 2 < n ? y[i + 2] : 0.0;
    const double y3 = i + 3 < n ? y[i + 3] : 0.0;
#pragma unroll
    for (int j = 0; j < MDOT_MAX; j++)
      if (j < K) {
        const float4 v = *reinterpret_cast<const float4*>(Vb + j * ld + i);
        acc[j] += (double)v.x * y0 + (double)v.y * y1 + (double)v.z * y2 + (double)v.w * y3;
      }
  }
#pragma unroll
  for (int j = 0; j < MDOT_MAX; j++) {
    if (j < K) {
      const double s = bsum(acc[j], sh);
      if (threadIdx.x == 0) partial[j * MDOT_BLOCKS + blockIdx.x] = s;
    }
  }
}

__global__ void __launch_bounds__(256) k_gs_update32(long long n, int K, const double* __restrict__ h,
                                                     const float* __restrict__ Vb, long long ld,
                                                     double* __restrict__ y,
                                                     double* __restrict__ partial)
{
  __shared__ double sh[4];
  double hr[MDOT_MAX];
#pragma unroll
  for (int j = 0; j < MDOT_MAX; j++) hr[j] = j < K ? -h[j] : 0.0;
  double ss = 0.0;
  const long long nq = (n + 3) / 4;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += (long long)gridDim.x * blockDim.x) {
    const long long i = 4 * q;
    double v0 = y[i];
    double v1 = i + 1 < n ? y[i + 1] : 0.0;
    double v2 = i + 2 < n ? y[i + 2] : 0.0;
    double v3 = i + 3 < n ? y[i + 3] : 0.0;
#pragma unroll
    for (int j = 0; j < MDOT_MAX; j++)
      if (j < K) {
        const float4 w = *reinterpret_cast<const float4*>(Vb + j * ld + i);
        v0 += hr[j] * (double)w.x;
        v1 += hr[j] * (double)w.y;
        v2 += hr[j] * (double)w.z;
        v3 += hr[j] * (double)w.w;
      }
    if (i + 1 >= n) v1 = 0.0;         // past the end: padding of V, not part of y
    if (i + 2 >= n) v2 = 0.0;
    if (i + 3 >= n) v3 = 0.0;
    y[i] = v0;
    if (i + 1 < n) y[i + 1] = v1;
    if (i + 2 < n) y[i + 2] = v2;
    if (i + 3 < n) y[i + 3] = v3;
    ss += v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3;
  }
  const double s = bsum(ss, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[k] = sum of the MDOT_BLOCKS partials of dot k; one workgroup per dot
__global__ void k_mfinish(const double* __restrict__ partial, double* __restrict__ out)
{
  __shared__ double sh[4];
  const int k = blockIdx.x;
  double v = 0.0;
  for (int i = threadIdx.x; i < MDOT_BLOCKS; i += blockDim.x) v += partial[k * MDOT_BLOCKS + i];
  const double s = bsum(v, sh);
  if (threadIdx.x == 0) out[k] = s;
}

static int vblocks(long long n) { return (int)std::min<long long>(4096, (n + 255) / 256); }

void dev_precond(Ctx& c, int prec, int kind, const double* in, double* out)
{
  if (kind == RMX_PREC_ILU0 || kind == RMX_PREC_SILU0) {
    ilu_apply(c, ilu_of(c, kind), prec, in, out, nullptr);
  } else if (kind == RMX_PREC_SILU0C) {
    coarse_apply(c, prec, in, out);
  } else if (kind == RMX_PREC_MPNN) {
    mpnn_apply(c, prec, in, out);
  } else if (kind == RMX_PREC_NONE) {
    k_copy<<<vblocks(c.n), 256, 0, c.stream>>>(c.n, in, out);
  } else {
    fail(RMX_ERR_ARG, "unknown preconditioner kind " + std::to_string(kind));
  }
}

static void precond(Ctx& c, int prec, int kind, const double* in, double* out)
{
  dev_precond(c, prec, kind, in, out);
}

void ensure_work(Ctx& c)
{
  for (int k = 0; k < 10; k++) c.w[k].alloc(c.n);
  c.red.alloc(64 + 40 * MDOT_BLOCKS);
  if (!c.h_red) HIPCHECK(hipHostMalloc(&c.h_red, 64 * sizeof(double), hipHostMallocDefault));
}

static double fetch(Ctx& c, const double* dev, int k = 1)
{
  HIPCHECK(hipMemcpyAsync(c.h_red, dev, sizeof(double) * k, hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  return c.h_red[0];
}

// BiCGStab on A(prec) x = b, x0 = 0.  Returns convergence; x gets the solution.
static bool bicgstab(Ctx& c, int prec, int pk, const double* b, double* x, int maxIter, double eps,
                     int& iters, double& accuracy)
{
  const long long n = c.n;
  double *r0 = c.w[0].p, *r = c.w[1].p, *s = c.w[2].p, *t = c.w[3].p, *v = c.w[4].p,
         *p = c.w[5].p, *tmp = c.w[6].p;
  double* sc = c.red.p;
  const int vb = vblocks(n);
  k_copy3<<<vb, 256, 0, c.stream>>>(n, b, r, r0, p, x);
  {
    const double* xs[1] = { r0 };
    const double* ys[1] = { r };
    dev_dots(c, 1, xs, ys, sc + S_RHO);
  }
  const double resid = std::sqrt(fetch(c, sc + S_RHO));
  double r0r0 = resid * resid;
  accuracy = 1.0;
  iters = 0;
  bool conv = false;
  if (!(resid > 0.0)) { iters = 0; accuracy = 0.0; return true; }
  int cur = 0;
  const bool trace = getenv("RMX_TRACE") != nullptr;
  while (!conv) {
    iters++;
    precond(c, prec, pk, p, tmp);
    dev_spmv(c, prec, tmp, v);
    {
      const double* xs[1] = { r0 };
      const double* ys[1] = { v };
      dev_dots(c, 1, xs, ys, sc + S_RV);
    }
    k_bicg_s<<<vb, 256, 0, c.stream>>>(n, sc, S_RHO + cur, r, v, s);
    precond(c, prec, pk, s, tmp);
    dev_spmv(c, prec, tmp, t);
    {
      const double* xs[2] = { t, t };
      const double* ys[2] = { s, t };
      dev_dots(c, 2, xs, ys, sc + S_TS);
    }
    k_bicg_xr<<<vb, 256, 0, c.stream>>>(n, sc, S_RHO + cur, p, s, t, x, r);
    {
      // rho_new into the other rho slot, (r,r) into S_RR
      const double* xs[2] = { r0, r };
      const double* ys[2] = { r, r };
      dev_dots(c, 1, xs, ys, sc + S_RHO + (cur ^ 1));
      dev_dots(c, 1, xs + 1, ys + 1, sc + S_RR);
    }
    k_bicg_p<<<vb, 256, 0, c.stream>>>(n, sc, S_RHO + cur, S_RHO + (cur ^ 1), r, v, p);
    HIPCHECK(hipGetLastError());
    cur ^= 1;
    fetch(c, sc, 7);
    const double rr = c.h_red[S_RR];
    if (trace)
      fprintf(stderr, "[rmx] bicgstab %d: rho %.6e -> %.6e (r0,v) %.6e (t,s) %.6e (t,t) %.6e (r,r) %.6e\n",
              iters, c.h_red[S_RHO + (cur ^ 1)], c.h_red[S_RHO + cur], c.h_red[S_RV], c.h_red[S_TS],
              c.h_red[S_TT], rr);
    const double relacc = std::sqrt(rr) / resid;
    accuracy = relacc;
    if (relacc < eps) conv = true;
    else if (iters >= maxIter || !std::isfinite(relacc)) break;
    else if (std::fabs(c.h_red[S_RHO + cur]) < 1.0e-10 * std::sqrt(r0r0 * rr)) {
      // breakdown: the shadow residual r0 has become orthogonal to r (the reference's
      // recurrences then divide by rho = 0); restart with r0 = p = r, rho = (r, r)
      k_copy<<<vb, 256, 0, c.stream>>>(n, r, r0);
      k_copy<<<vb, 256, 0, c.stream>>>(n, r, p);
      HIPCHECK(hipMemcpyAsync(sc + S_RHO + cur, sc + S_RR, sizeof(double), hipMemcpyDeviceToDevice,
                              c.stream));
      r0r0 = rr;
      if (trace) fprintf(stderr, "[rmx] bicgstab %d: restart\n", iters);
    }
  }
  // x = M^-1 x
  k_copy<<<vb, 256, 0, c.stream>>>(n, x, tmp);
  precond(c, prec, pk, tmp, x);
  return conv;
}

// Flexible GMRES(m) on A(prec) x = b, x0 = 0.  In an FP32 solve the Gram-Schmidt dots and
// updates read an FP32 copy of the basis (written by the normalisation, half the bytes);
// the preconditioner input, the new vector and the solution update stay FP64.
static bool fgmres(Ctx& c, int prec, int pk, const double* b, double* x, int m, int maxIter,
                   double eps, int& iters, double& accuracy)
{
  const long long n = c.n;
  if (c.basis_m != m || c.V_basis.n != (size_t)(m + 1) * n) {
    c.V_basis.alloc((size_t)(m + 1) * n);
    c.Z_basis.alloc((size_t)m * n);
    c.V32_basis.release();
    c.basis_m = m;
  }
  const bool b32 = prec == RMX_F32;
  const long long ld32 = (n + 63) / 64 * 64;      // FP32 rows padded for 16-byte loads
  if (b32 && c.V32_basis.n != (size_t)(m + 1) * ld32) {
    c.V32_basis.alloc((size_t)(m + 1) * ld32);
    HIPCHECK(hipMemsetAsync(c.V32_basis.p, 0, sizeof(float) * c.V32_basis.n, c.stream));
  }
  double* Vb = c.V_basis.p;
  float* V32 = b32 ? c.V32_basis.p : nullptr;
  double* Zb = c.Z_basis.p;
  double* sc = c.red.p;
  double* partial = c.red.p + 64;
  const int vb = vblocks(n);
  HIPCHECK(hipMemsetAsync(x, 0, sizeof(double) * n, c.stream));
  std::vector<std::vector<double>> hh(m, std::vector<double>(m + 1, 0.0));
  std::vector<double> cs(m + 1), sn(m + 1), rs(m + 1);
  const double EPSILON = 1.0e-20, EPSMAC = 1.0e-16;
  double eps1 = 0.0, resid = 0.0, ro = 0.0;
  accuracy = 1.0;
  iters = 0;
  bool conv = false;
  DBuf<double> d_rs;
  d_rs.alloc(m + 1);
  while (true) {
    // vv[0] = (b - A x) / ||b - A x||
    dev_resid_p(c, prec, x, b, Vb);
    {
      const double* xs[1] = { Vb };
      dev_dots(c, 1, xs, xs, sc + S_B2);
    }
    ro = std::sqrt(fetch(c, sc + S_B2));
    if (std::fabs(ro) <= EPSILON) { conv = true; break; }
    k_scal_dev<<<vb, 256, 0, c.stream>>>(n, nullptr, 0, 1.0 / ro, Vb, V32);
    if (iters == 0) { resid = ro; eps1 = eps * ro; }
    const double ro_cycle = ro;
    rs[0] = ro;
    int i = -1;
    while (true) {
      i++;
      iters++;
      const int i1 = i + 1;
      double* vi = Vb + (size_t)i * n;
      double* vi1 = Vb + (size_t)i1 * n;
      double* zi = Zb + (size_t)i * n;
      precond(c, prec, pk, vi, zi);
      dev_spmv(c, prec, zi, vi1);
      // classical Gram-Schmidt: all dots, then all updates
      const int kt = kt_begin(c, KT_GS, b32 ? 8.0 * i1 * ld32 + 24.0 * n : 16.0 * i1 * n + 40.0 * n);
      if (b32) {
        k_mdot32<<<MDOT_BLOCKS, 256, 0, c.stream>>>(n, i1, V32, ld32, vi1, partial);
        k_mfinish<<<i1, 256, 0, c.stream>>>(partial, sc + 16);
        k_gs_update32<<<MDOT_BLOCKS, 256, 0, c.stream>>>(n, i1, sc + 16, V32, ld32, vi1, partial);
        k_mfinish<<<1, 256, 0, c.stream>>>(partial, sc + S_B2);
      } else {
        k_mdot<<<MDOT_BLOCKS, 256, 0, c.stream>>>(n, i1, Vb, n, vi1, partial);
        k_mfinish<<<i1, 256, 0, c.stream>>>(partial, sc + 16);
        k_gs_update<<<vb, 256, 0, c.stream>>>(n, i1, sc + 16, Vb, n, vi1);
        const double* xs[1] = { vi1 };
        dev_dots(c, 1, xs, xs, sc + S_B2);
      }
      kt_end(c, kt);
      HIPCHECK(hipMemcpyAsync(c.h_red, sc, sizeof(double) * (16 + i1), hipMemcpyDeviceToHost, c.stream));
      HIPCHECK(hipStreamSynchronize(c.stream));
      for (int j = 0; j <= i; j++) hh[i][j] = c.h_red[16 + j];
      double tn = std::sqrt(c.h_red[S_B2]);
      hh[i][i1] = tn;
      if (std::fabs(tn) > EPSILON)
        k_scal_dev<<<vb, 256, 0, c.stream>>>(n, sc, S_B2, 1.0, vi1, V32 ? V32 + i1 * ld32 : nullptr);
      else if (V32)      // no scaling (breakdown): the FP32 copy of the vector as it is
        k_scal_dev<<<vb, 256, 0, c.stream>>>(n, nullptr, 0, 1.0, vi1, V32 + i1 * ld32);
      for (int k = 1; k <= i; k++) {
        const int k1 = k - 1;
        const double tt = hh[i][k1];
        hh[i][k1] = cs[k1] * tt + sn[k1] * hh[i][k];
        hh[i][k] = -sn[k1] * tt + cs[k1] * hh[i][k];
      }
      double gam = std::sqrt(hh[i][i] * hh[i][i] + hh[i][i1] * hh[i][i1]);
      if (std::fabs(gam) <= EPSILON) gam = EPSMAC;
      cs[i] = hh[i][i] / gam;
      sn[i] = hh[i][i1] / gam;
      rs[i1] = -sn[i] * rs[i];
      rs[i] = cs[i] * rs[i];
      hh[i][i] = cs[i] * hh[i][i] + sn[i] * hh[i][i1];
      ro = std::fabs(rs[i1]);
      accuracy = ro / resid;
      if (i + 1 >= m || ro <= eps1 || iters >= maxIter || !std::isfinite(ro)) break;
    }
    // back substitution
    rs[i] = rs[i] / hh[i][i];
    for (int ii = 2; ii <= i + 1; ii++) {
      const int k = i - ii + 1, k1 = k + 1;
      double tt = rs[k];
      for (int j = k1; j <= i; j++) tt -= hh[j][k] * rs[j];
      rs[k] = tt / hh[k][k];
    }
    HIPCHECK(hipMemcpyAsync(d_rs.p, rs.data(), sizeof(double) * (i + 1), hipMemcpyHostToDevice, c.stream));
    k_comb<<<vb, 256, 0, c.stream>>>(n, i + 1, d_rs.p, Zb, n, x);
    HIPCHECK(hipStreamSynchronize(c.stream));
    if (getenv("RMX_TRACE"))
      fprintf(stderr, "[rmx] fgmres(%d) %d: cycle end, residual estimate %.6e of the start\n", m, iters,
              ro / resid);
    if (ro <= eps1) { conv = true; break; }
    if (iters >= maxIter || !std::isfinite(ro)) { conv = false; break; }
    // a restart cycle that lowered the residual by less than a thousandth has stagnated
    // (restarted GMRES can stall on indefinite systems): end the solve unconverged
    if (ro > 0.999 * ro_cycle) { conv = false; break; }
  }
  return conv;
}

static bool krylov(Ctx& c, int method, int prec, int pk, const double* b, double* x, int maxIter,
                   double eps, int mkyrl, int& iters, double& acc)
{
  if (method == RMX_FGMRES) return fgmres(c, prec, pk, b, x, mkyrl, maxIter, eps, iters, acc);
  return bicgstab(c, prec, pk, b, x, maxIter, eps, iters, acc);
}

static double true_resid(Ctx& c, const double* x, double* r)
{
  dev_resid(c, x, c.B.p, r);
  const double* xs[2] = { r, c.B.p };
  dev_dots(c, 2, xs, xs, c.red.p + 8);
  HIPCHECK(hipMemcpyAsync(c.h_red, c.red.p + 8, sizeof(double) * 2, hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  const double nb = std::sqrt(c.h_red[1]);
  return nb > 0 ? std::sqrt(c.h_red[0]) / nb : 0.0;
}

static void check_opts(Ctx& c, const rmx_krylov_opts& o)
{
  if (o.method == RMX_FGMRES && (o.mkyrl < 1 || o.mkyrl > MDOT_MAX - 2))
    fail(RMX_ERR_ARG, "solve: FGMRES restart length must be 1..30");
  if (o.method != RMX_FGMRES && o.method != RMX_BICGSTAB) fail(RMX_ERR_ARG, "solve: unknown method");
  const int prec = o.prec == RMX_F32 ? RMX_F32 : RMX_F64;
  const int pk = o.precond;
  if (pk == RMX_PREC_ILU0 || pk == RMX_PREC_SILU0 || pk == RMX_PREC_SILU0C || pk == RMX_PREC_MPNN) {
    Ilu& I = ilu_of(c, pk);
    if (!((prec == RMX_F32) ? I.fac32 : I.fac64))
      fail(RMX_ERR_ARG, "solve: ILU(0) factors of the requested precision are missing");
  }
  if ((pk == RMX_PREC_SILU0C || pk == RMX_PREC_MPNN) && !(c.coarse.fac && c.coarse.fac_prec == prec))
    fail(RMX_ERR_ARG, "solve: coarse space not factored in the requested precision");
  if (pk == RMX_PREC_MPNN && !mpnn_ready(c)) fail(RMX_ERR_ARG, "solve: no learned preconditioner loaded");
  if ((prec == RMX_F32 && !c.have32) || (prec == RMX_F64 && !c.have64))
    fail(RMX_ERR_ARG, "solve: matrix of the requested precision is missing");
  if (o.refine && !c.have64) fail(RMX_ERR_ARG, "refinement needs the FP64 matrix");
}

// One solver of the chain on A X = B, from X = 0 or from the current X (continue).
static void solve_one(Ctx& c, const rmx_krylov_opts& o, bool from_zero, rmx_krylov_out& out)
{
  const int prec = o.prec == RMX_F32 ? RMX_F32 : RMX_F64;
  const int pk = o.precond;
  out = rmx_krylov_out{};
  if (!o.refine) {
    int it = 0; double acc = 1.0;
    if (from_zero) {
      const bool conv = krylov(c, o.method, prec, pk, c.B.p, c.X.p, o.max_iter, o.eps, o.mkyrl, it, acc);
      out.converged = conv;
      out.accuracy = acc;
    } else {
      // continue from X: the residual equation A d = B - A X, X += d, the accuracy relative
      // to ||B|| as a solve from zero reports it
      double* r = c.w[7].p;
      double* d = c.w[8].p;
      dev_resid_p(c, prec, c.X.p, c.B.p, r);
      const double* xs[2] = { r, c.B.p };
      dev_dots(c, 2, xs, xs, c.red.p + 8);
      HIPCHECK(hipMemcpyAsync(c.h_red, c.red.p + 8, sizeof(double) * 2, hipMemcpyDeviceToHost, c.stream));
      HIPCHECK(hipStreamSynchronize(c.stream));
      const double rel = c.h_red[1] > 0 ? std::sqrt(c.h_red[0] / c.h_red[1]) : 0.0;
      const bool conv = rel <= o.eps ||
                        krylov(c, o.method, prec, pk, r, d, o.max_iter, o.eps / std::max(rel, 1e-300),
                               o.mkyrl, it, acc);
      if (rel > o.eps) k_axpy<<<vblocks(c.n), 256, 0, c.stream>>>(c.n, 1.0, d, c.X.p);
      out.converged = conv;
      out.accuracy = rel > o.eps ? acc * rel : rel;
    }
    out.iters = it;
    out.true_resid = c.have64 ? true_resid(c, c.X.p, c.w[7].p) : -1.0;
    return;
  }
  // FP64 iterative refinement: r = B - A64 x; repeat: solve A_p d = r; x += d
  double* xr = c.X.p;
  double* r = c.w[7].p;
  double* d = c.w[8].p;
  double* xbest = c.w[9].p;                    // the iterate of the smallest true residual
  double tr = 1.0;
  if (from_zero) {
    HIPCHECK(hipMemsetAsync(xr, 0, sizeof(double) * c.n, c.stream));
    k_copy<<<vblocks(c.n), 256, 0, c.stream>>>(c.n, c.B.p, r);
  } else {
    tr = true_resid(c, xr, r);
  }
  int total = 0;
  const int sweeps_max = o.refine_max > 0 ? o.refine_max : 10;
  int sw = 0;
  int method = o.method, kind = pk, mk = o.mkyrl;
  bool bicg_tried = method == RMX_BICGSTAB;
  double best = tr;
  if (tr > o.refine_tol) {
    k_copy<<<vblocks(c.n), 256, 0, c.stream>>>(c.n, xr, xbest);
    for (; sw < sweeps_max; sw++) {
      int it = 0; double acc = 1.0;
      const double tr_prev = tr;
      const int method_prev = method, kind_prev = kind;
      // a sweep solves for the remaining factor only: the inner tolerance is relaxed to half
      // of what takes the true residual from tr to refine_tol, never tightened below eps
      const double eps_k = std::max(o.eps, 0.5 * o.refine_tol / tr);
      krylov(c, method, prec, kind, r, d, o.max_iter, std::min(eps_k, 0.5), mk, it, acc);
      total += it;
      out.accuracy = acc;
      k_axpy<<<vblocks(c.n), 256, 0, c.stream>>>(c.n, 1.0, d, xr);
      tr = true_resid(c, xr, r);
      if (getenv("RMX_DEBUG"))
        fprintf(stderr, "[rmx] refinement sweep %d: %d iterations (inner %.2e), FP64 residual %.3e\n",
                sw + 1, it, acc, tr);
      if (tr < best) {
        best = tr;
        k_copy<<<vblocks(c.n), 256, 0, c.stream>>>(c.n, xr, xbest);
      }
      if (tr <= o.refine_tol) { sw++; break; }
      const bool rose = !(tr < 10.0 * tr_prev);
      if (rose) {
        // the sweep raised the true residual tenfold or more (a diverging inner solve): back
        // to the best iterate, and the rest of the solve on FGMRES with the longest restart,
        // which minimises the residual over its Krylov space; when that has run, nothing is
        // left (a smaller rise is refined from, as any other sweep)
        k_copy<<<vblocks(c.n), 256, 0, c.stream>>>(c.n, xbest, xr);
        tr = true_resid(c, xr, r);
        if (kind == RMX_PREC_MPNN) {
          kind = RMX_PREC_SILU0C;
          out.fallback = std::max(out.fallback, 1);
        } else if (!(method == RMX_FGMRES && mk == MDOT_MAX - 2)) {
          method = RMX_FGMRES;
          mk = MDOT_MAX - 2;
          out.fallback = 2;
        } else {
          sw++;
          break;
        }
      } else if (!(tr < 0.1 * tr_prev) && (kind == RMX_PREC_MPNN || (method == RMX_FGMRES && !bicg_tried))) {
        // the sweep stalled (the learned preconditioner, or FGMRES stagnating): the two-level
        // preconditioner in place of the learned one, then BiCGStab in place of FGMRES
        if (kind == RMX_PREC_MPNN) {
          kind = RMX_PREC_SILU0C;
        } else {
          method = RMX_BICGSTAB;
          bicg_tried = true;
        }
        out.fallback = std::max(out.fallback, 1);
      }
      if (getenv("RMX_DEBUG") && (rose || method != method_prev || kind != kind_prev))
        fprintf(stderr, "[rmx] refinement sweep %s: %s(%d), preconditioner %d from here\n",
                rose ? "raised the residual" : "stalled", method == RMX_FGMRES ? "FGMRES" : "BiCGStab",
                mk, kind);
    }
    if (best < tr || !std::isfinite(tr)) {
      // later sweeps raised the residual: X is the best iterate
      k_copy<<<vblocks(c.n), 256, 0, c.stream>>>(c.n, xbest, xr);
      tr = best;
    }
  }
  out.iters = total; out.sweeps = sw; out.true_resid = tr;
  out.converged = tr <= o.refine_tol;
  if (out.converged && sw == 0) out.accuracy = tr;
}

void dev_solve(Ctx& c, const rmx_krylov_opts& o0, rmx_krylov_out& out)
{
  check_opts(c, o0);
  ensure_work(c);
  c.X.alloc(c.n);
  const auto t0 = std::chrono::steady_clock::now();
  const rmx_krylov_opts* o = &o0;
  bool from_zero = true;
  int iters = 0, sweeps = 0, fallback = 0, solvers = 0;
  rmx_krylov_out one{};
  for (;;) {
    solve_one(c, *o, from_zero, one);
    solvers++;
    iters += one.iters;
    sweeps += one.sweeps;
    fallback = std::max(fallback, one.fallback);
    if (one.converged || o->proceed == 0) break;
    if (o->proceed < 0) {
      // the reference sets X to zero and interrupts (Solve.cpp:345-353)
      HIPCHECK(hipMemsetAsync(c.X.p, 0, sizeof(double) * c.n, c.stream));
      one.interrupted = 1;
      break;
    }
    const rmx_krylov_opts* nx = o->next;
    if (!nx) fail(RMX_ERR_ARG, "solve: proceed > 0 without a next solver");
    if (solvers >= 16) fail(RMX_ERR_ARG, "solve: proceed chain longer than 16 solvers");
    check_opts(c, *nx);
    const double reached = o->refine ? one.true_resid : one.accuracy;
    if (!(nx->eps < reached)) break;             // Solve.cpp:337: next only when tighter
    from_zero = reached >= 1.0;                  // Solve.cpp:331: X = 0 when nothing was gained
    o = nx;
  }
  out = one;
  out.iters = iters;
  out.sweeps = sweeps;
  out.fallback = fallback;
  out.solvers = solvers;
  HIPCHECK(hipStreamSynchronize(c.stream));
  kt_collect(c);
  out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace rmx
