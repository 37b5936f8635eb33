// CSR kernels on gfx950: SpMV, residual, deterministic reductions, L2 scaling and
// precision conversion of the system matrix.
//
// Layout: one CSR per system (diagonal first in every row, columns in the reference's
// insertion order), FP64 and/or FP32 values.  All bandwidth-bound: SpMV moves
// z*(v+4) + 4(n+1) + 16n bytes (z non-zeros, v value bytes).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>

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

// ---------------------------------------------------------------------------------------
// FP32 SpMV over row groups.  G lanes per group, U columns per lane per pass: lane k reads
// a column's offset and x entry once and multiplies it into the partial sums of the group's
// rows (up to 4; the values of row r sit W entries after those of row r-1, so every row's
// loads are coalesced).  All loads of a pass are issued before the first is used: indices
// past the group (rows >= R, columns >= W) are clamped to valid entries of the group and
// their products zeroed by selects, so no load sits behind a branch.  The four partial sums
// are reduced together: two exchanges halve the rows a lane carries (4 -> 2 -> 1), then a
// plain shuffle tree finishes, so lane block r (G/4 lanes) holds row r's sum.

template <int G, int U, bool RESID>
__global__ void __launch_bounds__(256) k_spmv_grp(long long ng, const int* __restrict__ row0,
                                                  const int4* __restrict__ meta,
                                                  const uint16_t* __restrict__ off,
                                                  const float* __restrict__ A,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ y)
{
  static_assert(G >= 4 && G <= 64 && U >= 1, "lanes per group");
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long g = t / G;
  const int lane = (int)(t % G);
  if (g >= ng) return;               // whole G-lane segments leave together
  const int r0 = row0[g], R = row0[g + 1] - r0;
  const int4 m = meta[g];            // first value, first offset, smallest column, columns
  const int W = m.w, cb = m.z;
  const uint16_t* o = off + m.y;
  const float* a = A + m.x;
  const int last = R - 1;
  const int ro1 = min(1, last) * W, ro2 = min(2, last) * W, ro3 = min(3, last) * W;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int k0 = 0;
  do {                               // W >= 1: at least one pass

    int kk[U];
    float v0[U], v1[U], v2[U], v3[U];
    uint16_t oc[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      kk[u] = k0 + u * G + lane;
      const int k = min(kk[u], W - 1);
      oc[u] = o[k];
      v0[u] = a[k];
      v1[u] = a[ro1 + k];
      v2[u] = a[ro2 + k];
      v3[u] = a[ro3 + k];
    }
    double xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = x[cb + oc[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const double xe = kk[u] < W ? xv[u] : 0.0;
      s0 += (double)v0[u] * xe;
      s1 += (last >= 1 ? (double)v1[u] : 0.0) * xe;
      s2 += (last >= 2 ? (double)v2[u] : 0.0) * xe;
      s3 += (last >= 3 ? (double)v3[u] : 0.0) * xe;
    }
    k0 += G * U;
  } while (k0 < W);
  const int h = (lane / (G / 2)) & 1;          // upper half: rows 2, 3
  const double p0 = (h ? s2 : s0) + __shfl_xor(h ? s0 : s2, G / 2, G);
  const double p1 = (h ? s3 : s1) + __shfl_xor(h ? s1 : s3, G / 2, G);
  const int q = (lane / (G / 4)) & 1;          // odd row of the pair
  double v = (q ? p1 : p0) + __shfl_xor(q ? p0 : p1, G / 4, G);
#pragma unroll
  for (int w = G / 8; w > 0; w >>= 1) v += __shfl_xor(v, w, G);
  const int r = 2 * h + q;
  if ((lane & (G / 4 - 1)) == 0 && r < R) y[r0 + r] = RESID ? b[r0 + r] - v : v;
}

__global__ void k_gather_f32(long long n, const int* __restrict__ perm, const float* __restrict__ a,
                             float* __restrict__ out)
{
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    out[t] = a[perm[t]];
}

// values of the group layout current with A32 (gathered on the context stream)
static void spmv_groups_refresh(Ctx& c)
{
  SpmvGroups& S = c.sg;
  if (S.version == c.a32_version) return;
  k_gather_f32<<<4096, 256, 0, c.stream>>>(c.nnz, S.perm.p, c.A32.p, S.vals.p);
  HIPCHECK(hipGetLastError());
  S.version = c.a32_version;
}

template <int G, int U, bool RESID>
static void spmv_grp(Ctx& c, const double* x, const double* b, double* y)
{
  const SpmvGroups& S = c.sg;
  const int blocks = (int)((S.ng * G + 255) / 256);
  k_spmv_grp<G, U, RESID><<<blocks, 256, 0, c.stream>>>(S.ng, S.row0.p, S.meta.p, S.off.p, S.vals.p,
                                                       x, b, y);
}

template <bool RESID>
static void launch_spmv_grp(Ctx& c, const double* x, const double* b, double* y)
{
  spmv_groups_refresh(c);
  switch (c.sg.lanes) {
    case 4: spmv_grp<4, 8, RESID>(c, x, b, y); break;
    case 8: spmv_grp<8, 4, RESID>(c, x, b, y); break;
    case 16: spmv_grp<16, 2, RESID>(c, x, b, y); break;
    default: spmv_grp<32, 1, RESID>(c, x, b, y); break;
  }
  HIPCHECK(hipGetLastError());
}

// lanes per row of the CSR kernel: the power of two nearest a quarter of the mean row width,
// 4..32 (each lane then holds about four entries of its row)
static void choose_spmv_lanes(Ctx& c)
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

// Row groups of the FP32 layout from the host pattern: row i joins row i-1's group when
// both hold the same set of columns and the group has fewer than 4 rows.  Not built (CSR
// kernel) when a group's columns span 65536 or more, or the values do not fit int offsets.
void spmv_setup(Ctx& c, const int* rp, const int* cl)
{
  choose_spmv_lanes(c);
  SpmvGroups& S = c.sg;
  S.on = false;
  S.version = -1;
  const int n = c.n;
  const long long nnz = c.nnz;
  if (const char* e = getenv("RMX_SPMV_GROUPS"))
    if (atoi(e) == 0) return;
  if (n <= 0 || nnz >= (1LL << 31) - 1) return;
  // every row's positions in ascending column order
  std::vector<int> perm(nnz);
  parallel_for(n, [&](long long a, long long e) {
    for (long long i = a; i < e; i++) {
      int* p = perm.data() + rp[i];
      const int w = rp[i + 1] - rp[i];
      for (int k = 0; k < w; k++) p[k] = rp[i] + k;
      std::sort(p, p + w, [&](int u, int v) { return cl[u] < cl[v]; });
    }
  });
  std::vector<unsigned char> same(n, 0);
  parallel_for(n, [&](long long a, long long e) {
    for (long long i = std::max(1LL, a); i < e; i++) {
      const int w = rp[i + 1] - rp[i];
      if (w != rp[i] - rp[i - 1]) continue;
      bool eq = true;
      for (int k = 0; k < w && eq; k++) eq = cl[perm[rp[i] + k]] == cl[perm[rp[i - 1] + k]];
      same[i] = eq;
    }
  });
  std::vector<int> row0;
  std::vector<int4> meta;
  std::vector<uint16_t> off;
  row0.reserve(n / 2 + 1);
  for (int i = 0; i < n;) {
    int R = 1;
    while (R < 4 && i + R < n && same[i + R]) R++;
    const int w = rp[i + 1] - rp[i];
    const int lo = cl[perm[rp[i]]], hi = cl[perm[rp[i] + w - 1]];
    if (hi - lo >= 65536) return;
    row0.push_back(i);
    meta.push_back(make_int4(rp[i], (int)off.size(), lo, w));
    for (int k = 0; k < w; k++) off.push_back((uint16_t)(cl[perm[rp[i] + k]] - lo));
    if ((long long)off.size() >= (1LL << 31) - 1) return;
    i += R;
  }
  row0.push_back(n);
  hipStream_t s = c.stream;
  S.ng = (long long)meta.size();
  S.noff = (long long)off.size();
  S.row0.upload(row0, s); S.meta.upload(meta, s); S.off.upload(off, s); S.perm.upload(perm, s);
  S.vals.alloc(nnz);
  const double wm = (double)S.noff / (double)S.ng;
  S.lanes = wm <= 20.0 ? 4 : 8;      // 32 columns per pass covers most groups in one
  if (const char* e = getenv("RMX_SPMV_GG")) {
    const int v = atoi(e);
    if (v == 4 || v == 8 || v == 16 || v == 32) S.lanes = v;
  }
  HIPCHECK(hipStreamSynchronize(s));
  S.on = true;
}

double spmv_f32_bytes(const Ctx& c)
{
  const SpmvGroups& S = c.sg;
  if (!S.on) return (double)c.nnz * 8.0 + 4.0 * (c.n + 1) + 16.0 * c.n;
  // values, offsets, per group row0 and its int4 descriptor, x once, y
  return (double)c.nnz * 4.0 + 2.0 * S.noff + 20.0 * S.ng + 16.0 * c.n;
}

// algorithmic bytes of one product in the layout it runs on (+ 8n for b in residual form)
static double spmv_bytes(const Ctx& c, int prec, bool resid)
{
  double b;
  if (prec == RMX_F32 && c.sg.on) b = spmv_f32_bytes(c);
  else b = (double)c.nnz * (prec == RMX_F32 ? 8.0 : 12.0) + 4.0 * (c.n + 1) + 16.0 * c.n;
  return b + (resid ? 8.0 * c.n : 0.0);
}

void dev_spmv(Ctx& c, int prec, const double* x, double* y)
{
  const int kt = kt_begin(c, KT_SPMV, spmv_bytes(c, prec, false));
  if (prec == RMX_F32) {
    if (c.sg.on) launch_spmv_grp<false>(c, x, nullptr, y);
    else launch_spmv<float, false>(c, c.A32.p, x, nullptr, y);
  } else {
    launch_spmv<double, false>(c, c.A64.p, x, nullptr, y);
  }
  kt_end(c, kt);
}

void dev_resid_p(Ctx& c, int prec, const double* x, const double* b, double* r)
{
  const int kt = kt_begin(c, KT_SPMV, spmv_bytes(c, prec, true));
  if (prec == RMX_F32) {
    if (c.sg.on) launch_spmv_grp<true>(c, x, b, r);
    else launch_spmv<float, true>(c, c.A32.p, x, b, r);
  } else {
    launch_spmv<double, true>(c, c.A64.p, x, b, r);
  }
  kt_end(c, kt);
}

void dev_resid(Ctx& c, const double* x, const double* b, double* r)
{
  const int kt = kt_begin(c, KT_RESID64, spmv_bytes(c, RMX_F64, true));
  launch_spmv<double, true>(c, c.A64.p, x, b, r);
  kt_end(c, kt);
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
  else {
    k_scale_rows<float><<<2048, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.A32.p, c.B.p, c.red.p);
    c.a32_version++;
  }
  HIPCHECK(hipGetLastError());
  c.mat_version++;
  HIPCHECK(hipMemcpyAsync(c.h_red, c.red.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  return std::sqrt(c.h_red[0]);
}

// ---------------------------------------------------------------------------------------
// allocation-free scaling of a caller-held device system (rmx_scale_l2_noalloc)
//
// The reduction's partial sums and the squared norm live in `scratch`, a file-scope device
// array: one copy per GPU, held until the process ends, NOALLOC_DOUBLES doubles whatever
// the system's size (one partial sum per workgroup of the fixed RED_BLOCKS grid).  Host
// code never passes `scratch` itself to a kernel (in host code the name is only a
// stand-in): its address on each device is looked up once, with that device current, and
// kept in scratch_ptr.  Every call on one GPU uses that GPU's copy; nothing in it needs to
// start at zero (every partial sum and the norm are written before they are read), and
// calls that overlap on one GPU race on it.

constexpr int NOALLOC_DOUBLES = RED_BLOCKS + 8;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static double* scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

void noalloc_setup(int device)
{
  if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    fail(RMX_ERR_ARG, "no-allocation scratch: device index " + std::to_string(device) +
                      " outside 0.." + std::to_string(NOALLOC_MAX_DEVICES - 1));
  std::lock_guard<std::mutex> g(scratch_mu);
  if (scratch_ptr[device]) return;
  int cur = 0;
  HIPCHECK(hipGetDevice(&cur));
  HIPCHECK(hipSetDevice(device));
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  (void)hipSetDevice(cur);
  if (e != hipSuccess || !p)
    fail(RMX_ERR_HIP, std::string("no-allocation scratch: hipGetSymbolAddress: ") +
                      hipGetErrorString(e));
  scratch_ptr[device] = static_cast<double*>(p);
}

__global__ void k_sqrt_to(const double* __restrict__ nrm2, double* __restrict__ out)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = sqrt(nrm2[0]);
}

// The kernels and grids of dev_scale_l2 (so the results are bit-identical to rmx_scale_l2 on
// the same system), with the partial sums and the norm in the device's scratch copy.
void dev_scale_l2_noalloc(int device, int n, const int* rowptr, int prec, void* A, double* B,
                          double* scale, hipStream_t s)
{
  noalloc_setup(device);
  double* const partial = scratch_ptr[device];         // RED_BLOCKS partial sums
  double* const nrm2 = partial + RED_BLOCKS;           // ||B||^2
  DotArgs a;
  for (int k = 0; k < 8; k++) { a.x[k] = B; a.y[k] = B; }
  k_dots_n<<<RED_BLOCKS, 256, 0, s>>>(n, 1, a, partial);
  k_finish<<<1, 256, 0, s>>>(1, partial, nrm2);
  if (prec == RMX_F32)
    k_scale_rows<float><<<2048, 256, 0, s>>>(n, rowptr, static_cast<float*>(A), B, nrm2);
  else
    k_scale_rows<double><<<2048, 256, 0, s>>>(n, rowptr, static_cast<double*>(A), B, nrm2);
  k_sqrt_to<<<1, 64, 0, s>>>(nrm2, scale);
  HIPCHECK(hipGetLastError());
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
    c.a32_version++;
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
