// CSR kernels on gfx950: SpMV, residual, deterministic reductions, L2 scaling and
// precision conversion of the system matrix.
//
// Layout: one CSR per system (diagonal first in every row, columns in the reference's
// insertion order), FP64 and/or FP32 values.  All bandwidth-bound: SpMV moves
// z*(v+4) + 4(n+1) + 16n bytes (z non-zeros, v value bytes).
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "rmx_internal.h"

namespace rmx {

constexpr int RED_BLOCKS = 512;      // partial sums per reduction (fixed: deterministic)

// ---------------------------------------------------------------------------------------
// SpMV: G lanes per row (G chosen from the mean row width so a row's values and columns
// are read in one or two coalesced passes), shuffle reduction, FP64 accumulation.
// RESID: r = b - A x instead of y = A x.

template <class VT, int G, bool RESID>
__global__ void __launch_bounds__(256) k_spmv(int n, const int* __restrict__ rowptr,
                                              const int* __restrict__ cols,
                                              const VT* __restrict__ A,
                                              const double* __restrict__ x,
                                              const double* __restrict__ b,
                                              double* __restrict__ y)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int lane = (int)(t % G);
  if (row >= n) return;
  const int p0 = rowptr[row], p1 = rowptr[row + 1];
  // four entries per lane per pass, all loads issued before the gathers of x
  double s = 0.0;
  for (int p = p0 + lane; p < p1; p += 4 * G) {
    VT a[4];
    int cc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int q = p + u * G;
      a[u] = q < p1 ? A[q] : (VT)0;
      cc[u] = q < p1 ? cols[q] : -1;
    }
    double xv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) xv[u] = cc[u] >= 0 ? x[cc[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++) s += (double)a[u] * xv[u];
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (lane == 0) y[row] = RESID ? b[row] - s : s;
}

template <class VT, bool RESID>
static void launch_spmv(Ctx& c, const VT* A, const double* x, const double* b, double* y)
{
  const int G = c.spmv_g;
  const int blocks = (int)(((long long)c.n * G + 255) / 256);
  switch (G) {
    case 4: k_spmv<VT, 4, RESID><<<blocks, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, A, x, b, y); break;
    case 8: k_spmv<VT, 8, RESID><<<blocks, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, A, x, b, y); break;
    case 16: k_spmv<VT, 16, RESID><<<blocks, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, A, x, b, y); break;
    default: k_spmv<VT, 32, RESID><<<blocks, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, A, x, b, y); break;
  }
  HIPCHECK(hipGetLastError());
}

// lanes per row: the power of two nearest a quarter of the mean row width, 4..32 (each
// lane then holds about four entries of its row)
void choose_spmv_lanes(Ctx& c)
{
  const double w = c.n > 0 ? (double)c.nnz / c.n : 1.0;
  int g = 4;
  while (g < 32 && 4.0 * g * 1.41421356 < w) g *= 2;
  if (const char* e = getenv("RMX_SPMV_G")) {
    const int v = atoi(e);
    if (v == 4 || v == 8 || v == 16 || v == 32) g = v;
  }
  c.spmv_g = g;
}

void dev_spmv(Ctx& c, int prec, const double* x, double* y)
{
  if (prec == RMX_F32) launch_spmv<float, false>(c, c.A32.p, x, nullptr, y);
  else launch_spmv<double, false>(c, c.A64.p, x, nullptr, y);
}

void dev_resid_p(Ctx& c, int prec, const double* x, const double* b, double* r)
{
  if (prec == RMX_F32) launch_spmv<float, true>(c, c.A32.p, x, b, r);
  else launch_spmv<double, true>(c, c.A64.p, x, b, r);
}

void dev_resid(Ctx& c, const double* x, const double* b, double* r)
{
  launch_spmv<double, true>(c, c.A64.p, x, b, r);
}

// ---------------------------------------------------------------------------------------
// block reduction helpers

__device__ __forceinline__ double block_sum(double v, double* sh)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; i++) s += sh[i];
  }
  __syncthreads();
  return s;   // valid in thread 0
}

// partial[k*RED_BLOCKS + block] = sum_i x_k[i]*y_k[i] for K pairs
template <int K>
__global__ void __launch_bounds__(256) k_dots(int n, const double* const* __restrict__ xs,
                                              const double* const* __restrict__ ys,
                                              double* __restrict__ partial)
{
  __shared__ double sh[4];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] += xs[k][i] * ys[k][i];
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const double s = block_sum(acc[k], sh);
    if (threadIdx.x == 0) partial[k * RED_BLOCKS + blockIdx.x] = s;
  }
}

// out[k] = sum of partials (fixed order)
__global__ void k_finish(int K, const double* __restrict__ partial, double* __restrict__ out)
{
  __shared__ double sh[4];
  for (int k = 0; k < K; k++) {
    double v = 0.0;
    for (int i = threadIdx.x; i < RED_BLOCKS; i += blockDim.x) v += partial[k * RED_BLOCKS + i];
    const double s = block_sum(v, sh);
    if (threadIdx.x == 0) out[k] = s;
  }
}

struct DotArgs {
  const double* x[8];
  const double* y[8];
};

__global__ void __launch_bounds__(256) k_dots_n(int n, int K, DotArgs a, double* __restrict__ partial)
{
  __shared__ double sh[4];
  double acc[8];
  for (int k = 0; k < 8; k++) acc[k] = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    for (int k = 0; k < K; k++) acc[k] += a.x[k][i] * a.y[k][i];
  }
  for (int k = 0; k < K; k++) {
    const double s = block_sum(acc[k], sh);
    if (threadIdx.x == 0) partial[k * RED_BLOCKS + blockIdx.x] = s;
  }
}

// dots of up to 8 pairs into dst[0..K) (device), deterministic
void dev_dots(Ctx& c, int K, const double* const* x, const double* const* y, double* dst)
{
  DotArgs a;
  for (int k = 0; k < K; k++) { a.x[k] = x[k]; a.y[k] = y[k]; }
  for (int k = K; k < 8; k++) { a.x[k] = x[0]; a.y[k] = y[0]; }
  double* partial = c.red.p + 64;
  k_dots_n<<<RED_BLOCKS, 256, 0, c.stream>>>(c.n, K, a, partial);
  k_finish<<<1, 256, 0, c.stream>>>(K, partial, dst);
  HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// scaling: B /= ||B||, A /= ||B||  (CRSMAT::ScaleL2Norm)

template <class VT>
__global__ void k_scale_rows(int n, const int* __restrict__ rowptr, VT* __restrict__ A,
                             double* __restrict__ B, const double* __restrict__ nrm2)
{
  const double scale = sqrt(nrm2[0]);
  if (!(fabs(scale) > 1.0e-40)) return;
  const VT sv = (VT)scale;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    B[t] /= scale;
  }
  const long long nnz = rowptr[n];
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nnz;
       t += (long long)gridDim.x * blockDim.x)
    A[t] /= sv;
}

double dev_scale_l2(Ctx& c)
{
  const double* xs[1] = { c.B.p };
  dev_dots(c, 1, xs, xs, c.red.p);
  if (c.have64)
    k_scale_rows<double><<<2048, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.A64.p, c.B.p, c.red.p);
  else
    k_scale_rows<float><<<2048, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.A32.p, c.B.p, c.red.p);
  HIPCHECK(hipGetLastError());
  c.mat_version++;
  HIPCHECK(hipMemcpyAsync(c.h_red, c.red.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  return std::sqrt(c.h_red[0]);
}

__global__ void k_d2f(long long n, const double* __restrict__ a, float* __restrict__ b)
{
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    b[t] = (float)a[t];
}
__global__ void k_f2d(long long n, const float* __restrict__ a, double* __restrict__ b)
{
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    b[t] = (double)a[t];
}

void dev_convert(Ctx& c, int to_prec)
{
  if (to_prec == RMX_F32) {
    if (!c.have64) fail(RMX_ERR_ARG, "convert: no FP64 matrix");
    c.A32.alloc(c.nnz);
    k_d2f<<<4096, 256, 0, c.stream>>>(c.nnz, c.A64.p, c.A32.p);
    c.have32 = true;
  } else {
    if (!c.have32) fail(RMX_ERR_ARG, "convert: no FP32 matrix");
    c.A64.alloc(c.nnz);
    k_f2d<<<4096, 256, 0, c.stream>>>(c.nnz, c.A32.p, c.A64.p);
    c.have64 = true;
    c.mat_version++;
  }
  HIPCHECK(hipGetLastError());
}

}  // namespace rmx
