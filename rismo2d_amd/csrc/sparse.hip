// CSR kernels on gfx950: SpMV, deterministic reductions, L2 scaling, ILU(0)
// factorisation and triangular sweeps.
//
// Layout: one CSR per system (diagonal first in every row, columns in the reference's
// insertion order), FP64 and/or FP32 values.  The ILU(0) factors overwrite a copy of the
// values in the same layout (as PRECO_ILU0's shadow CRSMAT, CRSMat.cpp:110); after the
// factorisation they are split into a strictly-lower CSR, a strictly-upper CSR and the
// diagonal so each sweep streams only the half it needs.
//
// All bandwidth-bound: SpMV moves z*(v+4) + 4(n+1) + 16n bytes; a forward plus backward
// sweep moves about z*(v+4) + 32n bytes (z non-zeros, v value bytes).
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "rmx_internal.h"

namespace rmx {

constexpr int RED_BLOCKS = 512;      // partial sums per reduction (fixed: deterministic)

// ---------------------------------------------------------------------------------------
// SpMV: G lanes per row, shuffle reduction, FP64 accumulation

template <class VT, int G>
__global__ void __launch_bounds__(256) k_spmv(int n, const int* __restrict__ rowptr,
                                              const int* __restrict__ cols,
                                              const VT* __restrict__ A,
                                              const double* __restrict__ x,
                                              double* __restrict__ y)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int lane = (int)(t % G);
  if (row >= n) return;
  const int p0 = rowptr[row], p1 = rowptr[row + 1];
  double s = 0.0;
  for (int p = p0 + lane; p < p1; p += G) s += (double)A[p] * x[cols[p]];
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (lane == 0) y[row] = s;
}

// r = b - A x (FP64 true residual)
template <class VT, int G>
__global__ void __launch_bounds__(256) k_resid(int n, const int* __restrict__ rowptr,
                                               const int* __restrict__ cols,
                                               const VT* __restrict__ A,
                                               const double* __restrict__ x,
                                               const double* __restrict__ b,
                                               double* __restrict__ r)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / G);
  const int lane = (int)(t % G);
  if (row >= n) return;
  const int p0 = rowptr[row], p1 = rowptr[row + 1];
  double s = 0.0;
  for (int p = p0 + lane; p < p1; p += G) s += (double)A[p] * x[cols[p]];
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (lane == 0) r[row] = b[row] - s;
}

void dev_spmv(Ctx& c, int prec, const double* x, double* y)
{
  constexpr int G = 8;
  const int blocks = (int)(((long long)c.n * G + 255) / 256);
  if (prec == RMX_F32)
    k_spmv<float, G><<<blocks, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, c.A32.p, x, y);
  else
    k_spmv<double, G><<<blocks, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, c.A64.p, x, y);
  HIPCHECK(hipGetLastError());
}

void dev_resid_p(Ctx& c, int prec, const double* x, const double* b, double* r)
{
  constexpr int G = 8;
  const int blocks = (int)(((long long)c.n * G + 255) / 256);
  if (prec == RMX_F32)
    k_resid<float, G><<<blocks, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, c.A32.p, x, b, r);
  else
    k_resid<double, G><<<blocks, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, c.A64.p, x, b, r);
  HIPCHECK(hipGetLastError());
}

void dev_resid(Ctx& c, const double* x, const double* b, double* r)
{
  constexpr int G = 8;
  const int blocks = (int)(((long long)c.n * G + 255) / 256);
  k_resid<double, G><<<blocks, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, c.A64.p, x, b, r);
  HIPCHECK(hipGetLastError());
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
  }
  HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// ILU(0) factorisation (PRECO_ILU0::Factor, Preco_ilu0.cpp:64-186), row-wise: one wave
// per row, rows of one level in one launch.  Row i eliminates its lower entries in
// increasing column order k; each entry (i,j) receives its updates in increasing k, the
// same order as the reference's right-looking loop, so with FMA contraction off every
// factor entry is rounded exactly as the reference rounds it.  Multipliers are stored
// negated (L~ = -l), as the reference stores them.  Columns of row i are looked up by
// binary search in a per-row sorted copy held in LDS.

constexpr int FW = 128;   // maximum row width handled

template <class VT>
__global__ void __launch_bounds__(64) k_ilu_level(const int* __restrict__ rows, int cnt,
                                                   const int* __restrict__ rowptr,
                                                   const int* __restrict__ cols,
                                                   const int* __restrict__ srt_cols,
                                                   const unsigned char* __restrict__ srt_pos,
                                                   VT* __restrict__ LU, int* __restrict__ err)
{
#pragma clang fp contract(off)
  __shared__ VT s_val[1][FW];
  __shared__ int s_col[1][FW];
  __shared__ unsigned char s_pos[1][FW];
  const int wv = 0, lane = threadIdx.x;
  const int w = blockIdx.x;
  const bool active = w < cnt;
  const int i = active ? rows[w] : 0;
  const int p0 = active ? rowptr[i] : 0;
  const int wi = active ? rowptr[i + 1] - p0 : 0;
  for (int t = lane; t < wi; t += 64) {
    s_val[wv][t] = LU[p0 + t];
    s_col[wv][t] = srt_cols[p0 + t];
    s_pos[wv][t] = srt_pos[p0 + t];
  }
  __syncthreads();
  if (active) {
    for (int t = 0; t < wi; t++) {
      const int k = s_col[wv][t];
      if (k >= i) break;
      const int pik = s_pos[wv][t];
      const int q0 = rowptr[k], q1 = rowptr[k + 1];
      const VT ukk = LU[q0];
      const VT f = -s_val[wv][pik] / ukk;
      __syncthreads();
      if (lane == 0) s_val[wv][pik] = f;
      __syncthreads();
      for (int q = q0 + 1 + lane; q < q1; q += 64) {
        const int cc = cols[q];
        if (cc <= k) continue;
        int lo = 0, hi = wi - 1, pos = -1;
        while (lo <= hi) {
          const int mid = (lo + hi) >> 1;
          const int v = s_col[wv][mid];
          if (v == cc) { pos = s_pos[wv][mid]; break; }
          if (v < cc) lo = mid + 1; else hi = mid - 1;
        }
        if (pos >= 0) {
          const VT prod = f * LU[q];
          s_val[wv][pos] = s_val[wv][pos] + prod;
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  if (active) {
    for (int t = lane; t < wi; t += 64) LU[p0 + t] = s_val[wv][t];
    if (lane == 0 && fabs((double)s_val[wv][0]) < 1.0e-20) atomicOr(err, 1);
  }
}

// split the factor into strictly lower / strictly upper CSR and the diagonal
template <class VT>
__global__ void k_split(int n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                        const VT* __restrict__ LU, const int* __restrict__ lrp,
                        const int* __restrict__ urp, VT* __restrict__ lval,
                        VT* __restrict__ uval, VT* __restrict__ diag)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int pl = lrp[i], pu = urp[i];
  const int p0 = rowptr[i], p1 = rowptr[i + 1];
  diag[i] = LU[p0];
  for (int p = p0 + 1; p < p1; p++) {
    if (cols[p] < i) lval[pl++] = LU[p];
    else uval[pu++] = LU[p];
  }
}

// ---------------------------------------------------------------------------------------
// Triangular sweeps (PRECO_ILU0::Solve, Preco_ilu0.cpp:497-530, 958-980).
//   forward:  x_i = b_i + sum_{c<i} L~_ic x_c          (L~ stored negated)
//   backward: x_i = (x_i - sum_{c>i} U_ic x_c) / U_ii
// One workgroup walks the level sets in order; the rows of a level are spread over
// lane groups of G lanes; __syncthreads() separates levels (all waves of the group share
// the CU's L1, so stores of one level are visible to the next).  The row data of the
// next level is loaded before the barrier so HBM latency overlaps the current level.

template <class VT, int G, bool LOWER>
__global__ void __launch_bounds__(1024) k_sweep(int nlev, const int* __restrict__ lev_ptr,
                                                const int* __restrict__ lev_rows,
                                                const int* __restrict__ trp,
                                                const int* __restrict__ tcol,
                                                const VT* __restrict__ tval,
                                                const VT* __restrict__ diag,
                                                const double* __restrict__ b,
                                                double* __restrict__ x)
{
  const int gid = threadIdx.x / G, lane = threadIdx.x % G, ng = blockDim.x / G;
  for (int l = 0; l < nlev; l++) {
    const int a = lev_ptr[l], e = lev_ptr[l + 1];
    for (int r = a + gid; r < e; r += ng) {
      const int i = lev_rows[r];
      const int p0 = trp[i], p1 = trp[i + 1];
      double s = 0.0;
      for (int p = p0 + lane; p < p1; p += G) s += (double)tval[p] * x[tcol[p]];
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
      if (lane == 0) {
        if (LOWER) x[i] = b[i] + s;
        else x[i] = (x[i] - s) / (double)diag[i];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// Sweep, packed form.  Host set-up numbers each sweep's rows by processing position
// (level by level) and stores a level's off-diagonal entries in ELL layout (entry j of
// the level's row r at off + j*nr + r, padded to the level's widest row) with every column
// replaced by the position of that row.  Solved values are written twice: into an LDS
// ring (slot = position & (RING-1)) and into a level-ordered global copy xp.  A value at
// position p is still in the ring while the rows written since have not wrapped onto its
// slot, i.e. p >= P + nr - RING for the level starting at position P with nr rows;
// older values are read from xp, which is safe because a full barrier (global release /
// acquire) is taken whenever the start position crosses a multiple of FENCE, and an
// out-of-ring value is always older than the last full barrier (nr <= 1024,
// FENCE + 2*1024 <= RING).  All other levels are separated by an LDS-only barrier, so the
// next level's block, prefetched into registers while the current level computes, is
// not waited for.  The input (b, or the forward result for the backward sweep) is
// gathered into level order by a separate kernel.

constexpr int PW = 6;           // entries per lane held in registers (more are read on)
constexpr int SG = 8;           // lanes per row
constexpr int SKR = 2;          // row-rounds per level held in registers
constexpr int LPC = 1024;       // level offsets staged in LDS at a time
constexpr int FENCE = 2048;     // positions between full barriers

__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// x value at position c: the LDS ring when still there, else the level-ordered copy
__device__ __forceinline__ double ring_or_global(const double* ring, const double* xp, int c, int lim)
{
  double v;
  if (c >= lim) v = ring[c & (RING - 1)];
  else v = __builtin_nontemporal_load(xp + c);
  return v;
}

// G lanes per row: lane k takes entries k, k+G, ...; the first KR row-rounds of a level
// (KR * blockDim / G rows) with up to PW entries per lane are loaded into registers
// while the previous level finishes (one register set: the loads for level l+1 are
// issued right after level l has consumed its registers, so they fly across the
// barrier).  Wider levels and longer rows read the rest directly.
template <class VT, bool LOWER, int G, int KR>
__global__ void __launch_bounds__(1024) k_sweep_packed(int nlev, const int* __restrict__ lev_ptr,
                                                       const int* __restrict__ prs,
                                                       const int* __restrict__ prow,
                                                       const int* __restrict__ pcol,
                                                       const VT* __restrict__ pval,
                                                       const VT* __restrict__ pdiag,
                                                       const double* __restrict__ pin,
                                                       double* __restrict__ xp)
{
  __shared__ double ring[RING];
  __shared__ int s_lp[LPC + 2];
  const int t = threadIdx.x;
  const int grp = t / G, lane = t % G, ngr = blockDim.x / G;
  auto load_lp = [&](int l0) {
    for (int k = t; k < LPC + 2; k += blockDim.x) s_lp[k] = lev_ptr[min(l0 + k, nlev + 1)];
  };
  load_lp(0);
  __syncthreads();
  int lbase = 0;
  int cr[KR][PW];
  VT vr[KR][PW];
  int qar[KR], qbr[KR];
  double inr[KR];
  VT dgr[KR];
  int P = s_lp[0], nr = s_lp[1] - s_lp[0];
  auto fetch = [&](int P_, int nr_) {
#pragma unroll
    for (int k = 0; k < KR; k++) {
      const int g = grp + k * ngr;
      if (g < nr_) {
        const int r = P_ + g;
        inr[k] = pin[r];
        if (!LOWER) dgr[k] = pdiag[r];
        qar[k] = prs[r];
        qbr[k] = prs[r + 1];
#pragma unroll
        for (int j = 0; j < PW; j++) {
          const int q = qar[k] + lane + j * G;
          if (q < qbr[k]) { cr[k][j] = pcol[q]; vr[k][j] = pval[q]; }
          else { cr[k][j] = -1; vr[k][j] = (VT)0; }
        }
      }
    }
  };
  fetch(P, nr);
  for (int l = 0; l < nlev; l++) {
    const int lim = P + nr - RING;        // positions >= lim are still in the ring
    const int rounds = (nr + ngr - 1) / ngr;
    for (int k = 0; k < rounds; k++) {
      const int g = grp + k * ngr;
      const bool act = g < nr;
      double s = 0.0;
      double in = 0.0;
      VT dg = (VT)1;
      if (act) {
        if (k < KR) {
          in = inr[k < KR ? k : 0];
          dg = dgr[k < KR ? k : 0];
          bool old = false;
#pragma unroll
          for (int j = 0; j < PW; j++) {
            const int c = cr[k < KR ? k : 0][j];
            if (c >= 0) {
              old |= c < lim;
              s += (double)vr[k < KR ? k : 0][j] * ring[c & (RING - 1)];
            }
          }
          if (old) {                         // a dependency has left the ring: redo this row
            s = 0.0;
            for (int j = 0; j < PW; j++) {
              const int c = cr[k < KR ? k : 0][j];
              if (c >= 0) s += (double)vr[k < KR ? k : 0][j] * ring_or_global(ring, xp, c, lim);
            }
          }
          for (int q = qar[k < KR ? k : 0] + lane + PW * G; q < qbr[k < KR ? k : 0]; q += G)
            s += (double)pval[q] * ring_or_global(ring, xp, pcol[q], lim);
        } else {
          const int r = P + g;
          in = pin[r];
          if (!LOWER) dg = pdiag[r];
          for (int q = prs[r] + lane; q < prs[r + 1]; q += G)
            s += (double)pval[q] * ring_or_global(ring, xp, pcol[q], lim);
        }
      }
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
      if (act && lane == 0) {
        const double xi = LOWER ? in + s : (in - s) / (double)dg;
        ring[(P + g) & (RING - 1)] = xi;
        xp[P + g] = xi;
      }
    }
    const int Pn = P + nr;
    const bool full = Pn / FENCE != P / FENCE;
    if (l + 1 < nlev) {
      if (l + 2 - lbase > LPC) {            // next chunk of level offsets (rare)
        __syncthreads();
        lbase = l + 1;
        load_lp(lbase);
        __syncthreads();
      }
      P = s_lp[l + 1 - lbase];
      nr = s_lp[l + 2 - lbase] - P;
      fetch(P, nr);                         // next level's registers, in flight across the barrier
    }
    if (full) __syncthreads();
    else lds_barrier();
  }
}

__global__ void k_gather(int n, const int* __restrict__ idx, const double* __restrict__ src,
                         double* __restrict__ dst)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[r] = src[idx[r]];
}

__global__ void k_scatter(int n, const int* __restrict__ idx, const double* __restrict__ src,
                          double* __restrict__ dst)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[idx[r]] = src[r];
}

template <class VT>
__global__ void k_pack(long long ne, const int* __restrict__ psrc, const VT* __restrict__ LU,
                       VT* __restrict__ pval)
{
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ne;
       t += (long long)gridDim.x * blockDim.x) {
    const int s = psrc[t];
    pval[t] = s >= 0 ? LU[s] : (VT)0;
  }
}

template <class VT>
__global__ void k_pack_diag(int n, const int* __restrict__ prow, const int* __restrict__ rowptr,
                            const VT* __restrict__ LU, VT* __restrict__ pdiag)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) pdiag[r] = LU[rowptr[prow[r]]];
}

// Host: packed layout of one sweep: rows in processing order, each row's off-diagonal
// entries contiguous, columns replaced by positions.
static void build_pack(SweepPack& P, int n, const std::vector<int>& rp, const std::vector<int>& cl,
                       const std::vector<int>& lptr, const std::vector<int>& lrows, bool lower,
                       hipStream_t s)
{
  const int nlev = (int)lptr.size() - 1;
  P.nlev = nlev;
  P.h_lev = lptr;
  std::vector<int> pos(n);
  for (int r = 0; r < n; r++) pos[lrows[r]] = r;
  std::vector<int> prs(n + 1, 0);
  int maxnr = 0;
  for (int l = 0; l < nlev; l++) maxnr = std::max(maxnr, lptr[l + 1] - lptr[l]);
  for (int r = 0; r < n; r++) {
    const int i = lrows[r];
    int k = 0;
    for (int q = rp[i] + 1; q < rp[i + 1]; q++) k += lower ? (cl[q] < i) : (cl[q] > i);
    prs[r + 1] = prs[r] + k;
  }
  const long long ne = prs[n];
  P.ne = ne;
  P.max_nr = maxnr;
  std::vector<int> pcol(ne), psrc(ne);
  parallel_for(n, [&](long long a, long long b) {
    for (long long r = a; r < b; r++) {
      const int i = lrows[r];
      int j = prs[r];
      for (int q = rp[i] + 1; q < rp[i + 1]; q++)
        if (lower ? (cl[q] < i) : (cl[q] > i)) { pcol[j] = pos[cl[q]]; psrc[j] = q; j++; }
    }
  });
  P.ring_ok = maxnr <= 2048;          // FENCE + 2 * max rows per level <= RING
  if (getenv("RMX_DEBUG"))
    fprintf(stderr, "[rmx] %s sweep: %d levels, %.1f rows/level (max %d), %lld entries\n",
            lower ? "lower" : "upper", nlev, (double)n / nlev, maxnr, ne);
  std::vector<int> lp1(lptr);
  lp1.push_back(lptr.back());           // lev_ptr[l+2] is read for l = nlev-1
  P.lev_ptr.upload(lp1, s);
  P.prs.upload(prs, s);
  P.prow.upload(lrows, s);
  P.pcol.upload(pcol, s);
  P.psrc.upload(psrc, s);
  P.pin.alloc(n);
  P.xp.alloc(n);
}

template <class VT>
static const VT* pval_of(SweepPack& P);
template <>
const double* pval_of<double>(SweepPack& P) { return P.pval64.p; }
template <>
const float* pval_of<float>(SweepPack& P) { return P.pval32.p; }
template <class VT>
static const VT* pdiag_of(SweepPack& P);
template <>
const double* pdiag_of<double>(SweepPack& P) { return P.pdiag64.p; }
template <>
const float* pdiag_of<float>(SweepPack& P) { return P.pdiag32.p; }

static bool packed_ok(const Ctx& c) { return c.pk_lo.ring_ok && c.pk_up.ring_ok; }

template <class VT>
static void run_sweeps(Ctx& c, const VT* lval, const VT* uval, const VT* diag, const double* b,
                       double* x, hipEvent_t* ev)
{
  if (ev) HIPCHECK(hipEventRecord(ev[0], c.stream));
  if (packed_ok(c)) {
    SweepPack& L = c.pk_lo;
    SweepPack& U = c.pk_up;
    const int gb = (c.n + 255) / 256;
    k_gather<<<gb, 256, 0, c.stream>>>(c.n, L.prow.p, b, L.pin.p);
    k_sweep_packed<VT, true, SG, SKR><<<1, 1024, 0, c.stream>>>(L.nlev, L.lev_ptr.p, L.prs.p,
        L.prow.p, L.pcol.p, pval_of<VT>(L), nullptr, L.pin.p, L.xp.p);
    // forward result, from the forward level order straight into the backward level order
    k_gather<<<gb, 256, 0, c.stream>>>(c.n, c.fwd2bwd.p, L.xp.p, U.pin.p);
    if (ev) HIPCHECK(hipEventRecord(ev[1], c.stream));
    k_sweep_packed<VT, false, SG, SKR><<<1, 1024, 0, c.stream>>>(U.nlev, U.lev_ptr.p, U.prs.p,
        U.prow.p, U.pcol.p, pval_of<VT>(U), pdiag_of<VT>(U), U.pin.p, U.xp.p);
    k_scatter<<<gb, 256, 0, c.stream>>>(c.n, U.prow.p, U.xp.p, x);
  } else {
    constexpr int G = 4;
    k_sweep<VT, G, true><<<1, 1024, 0, c.stream>>>(c.lev_lo.nlev, c.lev_lo.d_ptr.p,
                                                    c.lev_lo.d_rows.p, c.lrp.p, c.lcol.p, lval,
                                                    diag, b, x);
    if (ev) HIPCHECK(hipEventRecord(ev[1], c.stream));
    k_sweep<VT, G, false><<<1, 1024, 0, c.stream>>>(c.lev_up.nlev, c.lev_up.d_ptr.p,
                                                     c.lev_up.d_rows.p, c.urp.p, c.ucol.p, uval,
                                                     diag, nullptr, x);
  }
  if (ev) HIPCHECK(hipEventRecord(ev[2], c.stream));
  HIPCHECK(hipGetLastError());
}

// Host: level sets of both sweeps, per-row sorted columns for the factorisation and the
// split lower / upper patterns.  Once per structure.
void setup_ilu_schedule(Ctx& c)
{
  const HostModel* hm = &c.hm;
  std::vector<int> hrp, hcols;
  if (c.csr_only) {
    hrp.resize(c.n + 1); hcols.resize(c.nnz);
    c.rowptr.download(hrp.data(), c.n + 1, c.stream);
    c.cols.download(hcols.data(), c.nnz, c.stream);
    HIPCHECK(hipStreamSynchronize(c.stream));
  }
  const std::vector<int>& rp = c.csr_only ? hrp : hm->rowptr;
  const std::vector<int>& cl = c.csr_only ? hcols : hm->cols;
  const int n = c.n;
  host_levels(n, rp, cl, true, c.lev_lo.ptr, c.lev_lo.rows);
  host_levels(n, rp, cl, false, c.lev_up.ptr, c.lev_up.rows);
  c.lev_lo.nlev = (int)c.lev_lo.ptr.size() - 1;
  c.lev_up.nlev = (int)c.lev_up.ptr.size() - 1;
  c.lev_lo.d_ptr.upload(c.lev_lo.ptr, c.stream); c.lev_lo.d_rows.upload(c.lev_lo.rows, c.stream);
  c.lev_up.d_ptr.upload(c.lev_up.ptr, c.stream); c.lev_up.d_rows.upload(c.lev_up.rows, c.stream);
  std::vector<int> scol(c.nnz);
  std::vector<unsigned char> spos(c.nnz);
  std::vector<int> lrp(n + 1, 0), urp(n + 1, 0);
  int maxw = 0;
  for (int i = 0; i < n; i++) {
    const int w = rp[i + 1] - rp[i];
    if (w > FW) fail(RMX_ERR_RANGE, "ILU(0): row wider than 128 columns");
    maxw = std::max(maxw, w);
    int nl = 0;
    for (int p = rp[i] + 1; p < rp[i + 1]; p++) nl += cl[p] < i;
    lrp[i + 1] = lrp[i] + nl;
    urp[i + 1] = urp[i] + (w - 1 - nl);
  }
  c.max_width = maxw;
  parallel_for(n, [&](long long a, long long b) {
    std::vector<std::pair<int, int>> tmp;
    for (long long i = a; i < b; i++) {
      tmp.clear();
      for (int p = rp[i]; p < rp[i + 1]; p++) tmp.emplace_back(cl[p], p - rp[i]);
      std::sort(tmp.begin(), tmp.end());
      for (size_t t = 0; t < tmp.size(); t++) {
        scol[rp[i] + t] = tmp[t].first;
        spos[rp[i] + t] = (unsigned char)tmp[t].second;
      }
    }
  });
  std::vector<int> lcol(lrp[n]), ucol(urp[n]);
  for (int i = 0; i < n; i++) {
    int pl = lrp[i], pu = urp[i];
    for (int p = rp[i] + 1; p < rp[i + 1]; p++) {
      if (cl[p] < i) lcol[pl++] = cl[p];
      else ucol[pu++] = cl[p];
    }
  }
  c.srt_cols.upload(scol, c.stream);
  c.srt_pos.upload(spos, c.stream);
  c.lrp.upload(lrp, c.stream); c.urp.upload(urp, c.stream);
  c.lcol.upload(lcol, c.stream); c.ucol.upload(ucol, c.stream);
  c.nlo = lrp[n]; c.nup = urp[n];
  build_pack(c.pk_lo, n, rp, cl, c.lev_lo.ptr, c.lev_lo.rows, true, c.stream);
  build_pack(c.pk_up, n, rp, cl, c.lev_up.ptr, c.lev_up.rows, false, c.stream);
  {
    // position in the forward order of the row at each backward position
    std::vector<int> fpos(n), f2b(n);
    for (int r = 0; r < n; r++) fpos[c.lev_lo.rows[r]] = r;
    for (int r = 0; r < n; r++) f2b[r] = fpos[c.lev_up.rows[r]];
    c.fwd2bwd.upload(f2b, c.stream);
  }
  HIPCHECK(hipStreamSynchronize(c.stream));
}

template <class VT>
static DBuf<VT>& pval_buf(SweepPack& P);
template <>
DBuf<double>& pval_buf<double>(SweepPack& P) { return P.pval64; }
template <>
DBuf<float>& pval_buf<float>(SweepPack& P) { return P.pval32; }
template <class VT>
static DBuf<VT>& pdiag_buf(SweepPack& P);
template <>
DBuf<double>& pdiag_buf<double>(SweepPack& P) { return P.pdiag64; }
template <>
DBuf<float>& pdiag_buf<float>(SweepPack& P) { return P.pdiag32; }

template <class VT>
static void factor_impl(Ctx& c, const VT* A, DBuf<VT>& LU, DBuf<VT>& lval, DBuf<VT>& uval,
                        DBuf<VT>& diag)
{
  LU.alloc(c.nnz);
  HIPCHECK(hipMemcpyAsync(LU.p, A, sizeof(VT) * c.nnz, hipMemcpyDeviceToDevice, c.stream));
  int* err = reinterpret_cast<int*>(c.red.p + 8);
  HIPCHECK(hipMemsetAsync(err, 0, sizeof(int), c.stream));
  for (int l = 0; l < c.lev_lo.nlev; l++) {
    const int a = c.lev_lo.ptr[l], cnt = c.lev_lo.ptr[l + 1] - a;
    k_ilu_level<VT><<<cnt, 64, 0, c.stream>>>(c.lev_lo.d_rows.p + a, cnt, c.rowptr.p, c.cols.p,
                                               c.srt_cols.p, c.srt_pos.p, LU.p, err);
  }
  HIPCHECK(hipGetLastError());
  lval.alloc(c.nlo); uval.alloc(c.nup); diag.alloc(c.n);
  k_split<VT><<<(c.n + 255) / 256, 256, 0, c.stream>>>(c.n, c.rowptr.p, c.cols.p, LU.p, c.lrp.p,
                                                       c.urp.p, lval.p, uval.p, diag.p);
  HIPCHECK(hipGetLastError());
  if (packed_ok(c)) {
    for (SweepPack* P : { &c.pk_lo, &c.pk_up }) {
      DBuf<VT>& pv = pval_buf<VT>(*P);
      pv.alloc(std::max<long long>(P->ne, 1));
      if (P->ne > 0) k_pack<VT><<<4096, 256, 0, c.stream>>>(P->ne, P->psrc.p, LU.p, pv.p);
    }
    DBuf<VT>& pd = pdiag_buf<VT>(c.pk_up);
    pd.alloc(c.n);
    k_pack_diag<VT><<<(c.n + 255) / 256, 256, 0, c.stream>>>(c.n, c.pk_up.prow.p, c.rowptr.p, LU.p, pd.p);
    HIPCHECK(hipGetLastError());
  }
  int h_err = 0;
  HIPCHECK(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  if (h_err) fail(RMX_ERR_PIVOT, "ILU(0): pivot below 1e-20 (row repair is not supported)");
}

void dev_ilu0_factor(Ctx& c, int prec)
{
  if (prec == RMX_F32) {
    if (!c.have32) fail(RMX_ERR_ARG, "ILU(0) FP32: no FP32 matrix (assemble or convert first)");
    factor_impl<float>(c, c.A32.p, c.L32, c.lval32, c.uval32, c.diag32);
    c.fac32 = true;
  } else {
    if (!c.have64) fail(RMX_ERR_ARG, "ILU(0) FP64: no FP64 matrix");
    factor_impl<double>(c, c.A64.p, c.L64, c.lval64, c.uval64, c.diag64);
    c.fac64 = true;
  }
}

void dev_ilu0_apply(Ctx& c, int prec, const double* b, double* x)
{
  if (prec == RMX_F32) {
    if (!c.fac32) fail(RMX_ERR_ARG, "ILU(0) FP32 not factored");
    run_sweeps<float>(c, c.lval32.p, c.uval32.p, c.diag32.p, b, x, nullptr);
  } else {
    if (!c.fac64) fail(RMX_ERR_ARG, "ILU(0) FP64 not factored");
    run_sweeps<double>(c, c.lval64.p, c.uval64.p, c.diag64.p, b, x, nullptr);
  }
}

// Average kernel durations (HIP events on the context stream) of SpMV and of the two
// sweeps, `reps` times each on the current system with a unit-ish vector.
void dev_profile(Ctx& c, int prec, int reps, rmx_kernel_times& out)
{
  hipEvent_t ev[3];
  for (auto& e : ev) HIPCHECK(hipEventCreate(&e));
  double* xv = c.w[0].p;
  double* yv = c.w[1].p;
  double t_spmv = 0, t_fwd = 0, t_bwd = 0;
  for (int r = 0; r < reps; r++) {
    HIPCHECK(hipEventRecord(ev[0], c.stream));
    dev_spmv(c, prec, c.B.p, yv);
    HIPCHECK(hipEventRecord(ev[1], c.stream));
    HIPCHECK(hipEventSynchronize(ev[1]));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    t_spmv += ms * 1e-3;
    if ((prec == RMX_F32 && c.fac32) || (prec != RMX_F32 && c.fac64)) {
      if (prec == RMX_F32) run_sweeps<float>(c, c.lval32.p, c.uval32.p, c.diag32.p, c.B.p, xv, ev);
      else run_sweeps<double>(c, c.lval64.p, c.uval64.p, c.diag64.p, c.B.p, xv, ev);
      HIPCHECK(hipEventSynchronize(ev[2]));
      float m1 = 0, m2 = 0;
      HIPCHECK(hipEventElapsedTime(&m1, ev[0], ev[1]));
      HIPCHECK(hipEventElapsedTime(&m2, ev[1], ev[2]));
      t_fwd += m1 * 1e-3;
      t_bwd += m2 * 1e-3;
    }
  }
  for (auto& e : ev) HIPCHECK(hipEventDestroy(e));
  out.spmv_s = t_spmv / reps;
  out.ilu_fwd_s = t_fwd / reps;
  out.ilu_bwd_s = t_bwd / reps;
  out.ilu_apply_s = out.ilu_fwd_s + out.ilu_bwd_s;
  out.n_spmv = reps;
  out.n_ilu = reps;
}

}  // namespace rmx
