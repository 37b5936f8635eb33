// Two-level preconditioner: strip ILU(0) sweeps with a coarse-space correction.
//
// The block-Jacobi strip sweeps only couple equations inside one block, so on a long
// channel information crosses the domain one block per Krylov iteration.  A coarse space
// with one unknown per aggregate (strip x coarse segment x variable U, V or H) carries the
// long-range part: with P the piecewise-constant prolongation and R = P^T,
//
//     z1 = P (R A P)^-1 R r,    z = z1 + M^-1 (r - A z1)        (M: strip ILU(0))
//
// i.e. a multiplicative coarse correction followed by the strip sweeps.  A z1 = (A P) zc
// with A P kept per row (a few aggregates per row), so the correction costs a pass over A P
// instead of a second SpMV.  R A P is summed on
// the device in a fixed order (every coarse entry is the sum of the matrix entries whose row
// and column aggregates match, listed by a permutation built once per pattern).  Aggregates
// are numbered strip by strip, so R A P is banded; its LU factorisation with partial
// pivoting (the layout and steps of LAPACK's band LU, gbtf2 / gbtrs) and the per-apply
// solve of a few thousand unknowns run on the host between two small copies.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <atomic>

#include "rmx_internal.h"

namespace rmx {

// cval[s] = sum over k in [slot_ptr[s], slot_ptr[s+1]) of A[perm[k]]: one wave per slot,
// lane-strided partial sums and a fixed shuffle tree (deterministic)
template <class VT>
__global__ void __launch_bounds__(256) k_coarse_vals(long long nslot, const long long* __restrict__ slot_ptr,
                                                     const int* __restrict__ perm,
                                                     const VT* __restrict__ A,
                                                     double* __restrict__ cval)
{
  const long long s = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (s >= nslot) return;
  double v = 0.0;
  for (long long k = slot_ptr[s] + lane; k < slot_ptr[s + 1]; k += 64) v += (double)A[perm[k]];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) cval[s] = v;
}

// rc[I] = sum of r over the rows of aggregate I (ascending rows, fixed order)
__global__ void __launch_bounds__(256) k_restrict(int nc, const int* __restrict__ agg_ptr,
                                                  const int* __restrict__ agg_rows,
                                                  const double* __restrict__ r,
                                                  double* __restrict__ rc)
{
  const int I = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (I >= nc) return;
  double v = 0.0;
  for (int k = agg_ptr[I] + lane; k < agg_ptr[I + 1]; k += 64) v += r[agg_rows[k]];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) rc[I] = v;
}

// apv[k] = sum of A over the entries of (row, aggregate) pair k (fixed order)
template <class VT>
__global__ void k_ap_vals(long long nap, const int* __restrict__ src_ptr, const int* __restrict__ src,
                          const VT* __restrict__ A, double* __restrict__ apv)
{
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nap) return;
  double v = 0.0;
  for (int q = src_ptr[k]; q < src_ptr[k + 1]; q++) v += (double)A[src[q]];
  apv[k] = v;
}

// t = r - (A P) zc
__global__ void k_ap_resid(int n, const int* __restrict__ ap_ptr, const int* __restrict__ ap_agg,
                           const double* __restrict__ apv, const double* __restrict__ zc,
                           const double* __restrict__ r, double* __restrict__ t)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = ap_ptr[i]; k < ap_ptr[i + 1]; k++) s += apv[k] * zc[ap_agg[k]];
  t[i] = r[i] - s;
}

// z += P zc
__global__ void k_add_prolong(int n, const int* __restrict__ agg, const double* __restrict__ zc,
                              double* __restrict__ z)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] += zc[agg[i]];
}

// ---------------------------------------------------------------------------------------
// host: banded LU with partial pivoting.  Column-major band storage with ldab = 2kl+ku+1;
// entry (i, j) at band[kl + ku + i - j + j*ldab]; the top kl rows hold the fill of the
// row interchanges.

static inline double& bel(std::vector<double>& ab, int ldab, int kv, int i, int j)
{
  return ab[(size_t)(kv + i - j) + (size_t)j * ldab];
}

static bool band_lu(int n, int kl, int ku, std::vector<double>& ab, std::vector<int>& ipiv)
{
  const int ldab = 2 * kl + ku + 1, kv = ku + kl;
  ipiv.assign(n, 0);
  int ju = 0;
  for (int j = 0; j < n; j++) {
    const int km = std::min(kl, n - 1 - j);
    int jp = 0;
    double best = std::fabs(bel(ab, ldab, kv, j, j));
    for (int t = 1; t <= km; t++) {
      const double v = std::fabs(bel(ab, ldab, kv, j + t, j));
      if (v > best) { best = v; jp = t; }
    }
    ipiv[j] = j + jp;
    if (!(best > 0.0)) return false;
    ju = std::max(ju, std::min(j + ku + jp, n - 1));
    if (jp != 0)
      for (int cc = j; cc <= ju; cc++) std::swap(bel(ab, ldab, kv, j, cc), bel(ab, ldab, kv, j + jp, cc));
    if (km > 0) {
      const double piv = bel(ab, ldab, kv, j, j);
      for (int t = 1; t <= km; t++) bel(ab, ldab, kv, j + t, j) /= piv;
      for (int cc = j + 1; cc <= ju; cc++) {
        const double u = bel(ab, ldab, kv, j, cc);
        if (u == 0.0) continue;
        for (int t = 1; t <= km; t++) bel(ab, ldab, kv, j + t, cc) -= bel(ab, ldab, kv, j + t, j) * u;
      }
    }
  }
  return true;
}

static void band_solve(int n, int kl, int ku, std::vector<double>& ab, const std::vector<int>& ipiv,
                       double* b)
{
  const int ldab = 2 * kl + ku + 1, kv = ku + kl;
  for (int j = 0; j < n - 1; j++) {
    const int km = std::min(kl, n - 1 - j);
    const int l = ipiv[j];
    if (l != j) std::swap(b[l], b[j]);
    const double bj = b[j];
    for (int t = 1; t <= km; t++) b[j + t] -= bel(ab, ldab, kv, j + t, j) * bj;
  }
  for (int j = n - 1; j >= 0; j--) {
    b[j] /= bel(ab, ldab, kv, j, j);
    const double bj = b[j];
    for (int i = std::max(0, j - kv); i < j; i++) b[i] -= bel(ab, ldab, kv, i, j) * bj;
  }
}

// ---------------------------------------------------------------------------------------
// block cyclic reduction of the coarse system (block tridiagonal in strips)
//
// Blocks i = 0..nb-1 (strip i's aggregates, padded to bs with identity rows), A_i the
// diagonal block, B_i the coupling to block i - s and C_i to block i + s at level l
// (s = 2^l).  Level l eliminates the odd multiples j of s:
//   kept i (multiples of 2s):  P_i = B_i A_{i-s}^-1,  Q_i = C_i A_{i+s}^-1,
//     A_i <- A_i - P_i C_{i-s} - Q_i B_{i+s},  B_i <- -P_i B_{i-s},  C_i <- -Q_i C_{i+s};
//   apply, reduction:        w_i <- w_i - P_i w_{i-s} - Q_i w_{i+s};
//   apply, back-substitution: w_j <- A_j^-1 (w_j - B_j w_{j-s} - C_j w_{j+s}).
// After the last level block 0 alone remains: w_0 <- A_0^-1 w_0.  Matrices are stored
// column-major (entry (r, c) at c * bs + r), so row r of a matrix-vector product reads
// consecutive entries across the lanes of a wave.

// w[i * bs + r] <- rc[blk_start[i] + r] (0 on padding)
__global__ void k_cr_in(int nb, int bs, const int* __restrict__ blk_start,
                        const double* __restrict__ rc, double* __restrict__ w)
{
  const int i = blockIdx.x, r = threadIdx.x;
  if (i >= nb || r >= bs) return;
  const int a = blk_start[i], m = blk_start[i + 1] - a;
  w[(long long)i * bs + r] = r < m ? rc[a + r] : 0.0;
}

__global__ void k_cr_out(int nb, int bs, const int* __restrict__ blk_start,
                         const double* __restrict__ w, double* __restrict__ zc)
{
  const int i = blockIdx.x, r = threadIdx.x;
  if (i >= nb || r >= bs) return;
  const int a = blk_start[i], m = blk_start[i + 1] - a;
  if (r < m) zc[a + r] = w[(long long)i * bs + r];
}

constexpr int CR_MAX = 128;       // rows of a coarse block (a strip's aggregates)

// one workgroup (a thread per block row) per kept block k (i = 2 s k) of level s:
// w_i -= P_i w_{i-s} + Q_i w_{i+s}
__global__ void __launch_bounds__(CR_MAX) k_cr_fwd(int nb, int bs, int s,
                                                   const double* __restrict__ M,
                                                   double* __restrict__ w)
{
  __shared__ double lo[CR_MAX], hi[CR_MAX];
  const int k = blockIdx.x, r = threadIdx.x;
  const int i = 2 * s * k;
  if (i >= nb) return;
  const bool hl = i - s >= 0, hh = i + s < nb;
  if (r < bs) {
    lo[r] = hl ? w[(long long)(i - s) * bs + r] : 0.0;
    hi[r] = hh ? w[(long long)(i + s) * bs + r] : 0.0;
  }
  __syncthreads();
  if (r >= bs) return;
  const long long b2 = (long long)bs * bs;
  const double* P = M + 2 * b2 * k;
  const double* Q = P + b2;
  double acc = 0.0;
  for (int c = 0; c < bs; c++) acc += P[c * bs + r] * lo[c] + Q[c * bs + r] * hi[c];
  w[(long long)i * bs + r] -= acc;
}

// one workgroup per eliminated block k (j = s + 2 s k):
// w_j <- A_j^-1 (w_j - B_j w_{j-s} - C_j w_{j+s}); with s = 0 the top block, w_0 <- A^-1 w_0
__global__ void __launch_bounds__(CR_MAX) k_cr_bwd(int nb, int bs, int s,
                                                   const double* __restrict__ M,
                                                   double* __restrict__ w)
{
  __shared__ double lo[CR_MAX], hi[CR_MAX], t[CR_MAX];
  const int k = blockIdx.x, r = threadIdx.x;
  const int j = s + 2 * s * k;
  if (j >= nb) return;
  const long long b2 = (long long)bs * bs;
  const double* Ai = M + (s ? 3 * b2 * k : 0);
  double v = 0.0;
  if (s) {
    const bool hl = j - s >= 0, hh = j + s < nb;
    if (r < bs) {
      lo[r] = hl ? w[(long long)(j - s) * bs + r] : 0.0;
      hi[r] = hh ? w[(long long)(j + s) * bs + r] : 0.0;
    }
    __syncthreads();
    if (r < bs) {
      const double* B = Ai + b2;
      const double* C = B + b2;
      v = w[(long long)j * bs + r];
      for (int c = 0; c < bs; c++) v -= B[c * bs + r] * lo[c] + C[c * bs + r] * hi[c];
    }
  } else if (r < bs) {
    v = w[(long long)j * bs + r];
  }
  if (r < bs) t[r] = v;
  __syncthreads();
  if (r >= bs) return;
  double x = 0.0;
  for (int c = 0; c < bs; c++) x += Ai[c * bs + r] * t[c];
  w[(long long)j * bs + r] = x;
}

// ---- factorisation of the reduction on the device (column-major bs x bs blocks) ----

// blocks[slot_dst[k]] = cval[k]: the level-0 blocks B (i, i-1), A (i, i), C (i, i+1)
__global__ void k_cr_scatter(long long nslot, const long long* __restrict__ dst,
                             const double* __restrict__ cval, double* __restrict__ blocks)
{
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nslot) blocks[dst[k]] = cval[k];
}

// rows past a block's aggregates: identity in A
__global__ void k_cr_pad(int nb, int bs, const int* __restrict__ blk_start, double* __restrict__ Am)
{
  const int i = blockIdx.x, r = threadIdx.x;
  if (i >= nb || r >= bs) return;
  if (r >= blk_start[i + 1] - blk_start[i]) Am[(long long)i * bs * bs + (long long)r * bs + r] = 1.0;
}

// One workgroup inverts an n x n block held in LDS (column-major) by Gauss-Jordan with
// partial pivoting; returns false (same in every thread) for a pivot below 1e-300.
__device__ bool lds_inverse(double* m, int n, int* perm, double* red_v, int* red_i, double* fcol)
{
  const int t = threadIdx.x, nt = blockDim.x;
  for (int c = 0; c < n; c++) {
    // pivot: the largest |m(r, c)|, r >= c (first one on ties)
    double best = -1.0;
    int bi = c;
    for (int r = c + t; r < n; r += nt) {
      const double v = fabs(m[c * n + r]);
      if (v > best) { best = v; bi = r; }
    }
    red_v[t] = best; red_i[t] = bi;
    __syncthreads();
    for (int o = nt / 2; o > 0; o >>= 1) {
      if (t < o) {
        const double v = red_v[t + o];
        const int j = red_i[t + o];
        if (v > red_v[t] || (v == red_v[t] && j < red_i[t])) { red_v[t] = v; red_i[t] = j; }
      }
      __syncthreads();
    }
    const int p = red_i[0];
    const bool ok = red_v[0] > 1.0e-300;
    __syncthreads();
    if (!ok) return false;
    if (t == 0) perm[c] = p;
    if (p != c)
      for (int k = t; k < n; k += nt) {
        const double a = m[k * n + c];
        m[k * n + c] = m[k * n + p];
        m[k * n + p] = a;
      }
    __syncthreads();
    const double piv = m[c * n + c];
    __syncthreads();
    // row c scaled (its pivot column becomes 1 / piv), then every other row eliminated with
    // the factors of column c read before any is overwritten
    for (int k = t; k < n; k += nt) m[k * n + c] = (k == c ? 1.0 : m[k * n + c]) / piv;
    for (int r = t; r < n; r += nt) fcol[r] = m[c * n + r];
    __syncthreads();
    for (int idx = t; idx < n * n; idx += nt) {
      const int r = idx % n, k = idx / n;
      if (r == c) continue;
      const double v = k == c ? 0.0 : m[k * n + r];
      m[k * n + r] = v - fcol[r] * m[k * n + c];
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, in reverse order
  for (int c = n - 1; c >= 0; c--) {
    const int p = perm[c];
    if (p != c)
      for (int r = t; r < n; r += nt) {
        const double a = m[c * n + r];
        m[c * n + r] = m[p * n + r];
        m[p * n + r] = a;
      }
    __syncthreads();
  }
  return true;
}

// Eliminated blocks of level s (j = s + 2 s k; s = 0: the top block): A_j^-1 (also kept in
// inv for the level's kept blocks), B_j and C_j into the level's storage E.
__global__ void __launch_bounds__(256) k_cr_inv(int nb, int bs, int s, const double* __restrict__ Am,
                                                const double* __restrict__ Bm,
                                                const double* __restrict__ Cm,
                                                double* __restrict__ E, double* __restrict__ inv,
                                                int* __restrict__ bad)
{
  extern __shared__ double sm[];
  __shared__ double red_v[256], fcol[CR_MAX];
  __shared__ int red_i[256], perm[CR_MAX];
  const int k = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int j = s ? s + 2 * s * k : 0;
  if (j >= nb) return;
  const long long b2 = (long long)bs * bs;
  for (int e = t; e < b2; e += nt) sm[e] = Am[j * b2 + e];
  __syncthreads();
  const bool ok = lds_inverse(sm, bs, perm, red_v, red_i, fcol);
  if (!ok) {
    if (t == 0) atomicOr(bad, 1);
    return;
  }
  double* Ai = E + (s ? 3 * b2 * k : 0);
  for (int e = t; e < b2; e += nt) {
    Ai[e] = sm[e];
    inv[j * b2 + e] = sm[e];
    if (s) {
      Ai[b2 + e] = Bm[j * b2 + e];
      Ai[2 * b2 + e] = Cm[j * b2 + e];
    }
  }
}

// out (+)= alpha a b, column-major bs x bs, by one workgroup
__device__ void wg_mm(int n, const double* a, const double* b, double alpha, bool acc, double* out)
{
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    const int r = idx % n, c = idx / n;
    double v = 0.0;
    for (int k = 0; k < n; k++) v += a[k * n + r] * b[c * n + k];
    out[idx] = (acc ? out[idx] : 0.0) + alpha * v;
  }
  __syncthreads();
}

// Kept blocks of level s (i = 2 s k): P = B_i A_{i-s}^-1, Q = C_i A_{i+s}^-1 into the
// level's storage K; A_i -= P C_{i-s} + Q B_{i+s}; B_i = -P B_{i-s}; C_i = -Q C_{i+s}.
// tmp: per kept block two scratch blocks.
__global__ void __launch_bounds__(256) k_cr_keep(int nb, int bs, int s, double* __restrict__ Am,
                                                 double* __restrict__ Bm, double* __restrict__ Cm,
                                                 const double* __restrict__ inv,
                                                 double* __restrict__ K, double* __restrict__ tmp)
{
  const int k = blockIdx.x, i = 2 * s * k;
  if (i >= nb) return;
  const long long b2 = (long long)bs * bs;
  double* P = K + 2 * b2 * k;
  double* Q = P + b2;
  double* nB = tmp + 2 * b2 * k;
  double* nC = nB + b2;
  for (int e = threadIdx.x; e < b2; e += blockDim.x) { P[e] = 0.0; Q[e] = 0.0; nB[e] = 0.0; nC[e] = 0.0; }
  __syncthreads();
  if (i - s >= 0) {
    wg_mm(bs, Bm + i * b2, inv + (i - s) * b2, 1.0, false, P);
    wg_mm(bs, P, Cm + (i - s) * b2, -1.0, true, Am + i * b2);
    wg_mm(bs, P, Bm + (i - s) * b2, -1.0, false, nB);
  }
  if (i + s < nb) {
    wg_mm(bs, Cm + i * b2, inv + (i + s) * b2, 1.0, false, Q);
    wg_mm(bs, Q, Bm + (i + s) * b2, -1.0, true, Am + i * b2);
    wg_mm(bs, Q, Cm + (i + s) * b2, -1.0, false, nC);
  }
  for (int e = threadIdx.x; e < b2; e += blockDim.x) {
    Bm[i * b2 + e] = nB[e];
    Cm[i * b2 + e] = nC[e];
  }
}

// host dense helpers on column-major bs x bs blocks
static bool dense_inverse(int n, const double* a, double* inv)
{
  std::vector<double> m(a, a + (size_t)n * n);
  std::vector<int> piv(n);
  for (int c = 0; c < n; c++) {                 // LU with partial pivoting, column c
    int p = c;
    double best = std::fabs(m[(size_t)c * n + c]);
    for (int r = c + 1; r < n; r++)
      if (std::fabs(m[(size_t)c * n + r]) > best) { best = std::fabs(m[(size_t)c * n + r]); p = r; }
    if (!(best > 1.0e-300)) return false;
    piv[c] = p;
    if (p != c)
      for (int k = 0; k < n; k++) std::swap(m[(size_t)k * n + c], m[(size_t)k * n + p]);
    const double d = m[(size_t)c * n + c];
    for (int r = c + 1; r < n; r++) m[(size_t)c * n + r] /= d;
    for (int k = c + 1; k < n; k++) {
      const double u = m[(size_t)k * n + c];
      if (u == 0.0) continue;
      for (int r = c + 1; r < n; r++) m[(size_t)k * n + r] -= m[(size_t)c * n + r] * u;
    }
  }
  for (int e = 0; e < n; e++) {                 // inverse column e: solve with unit vector
    double* x = inv + (size_t)e * n;
    for (int r = 0; r < n; r++) x[r] = r == e ? 1.0 : 0.0;
    for (int c = 0; c < n; c++) {
      if (piv[c] != c) std::swap(x[c], x[piv[c]]);
    }
    for (int c = 0; c < n; c++)
      for (int r = c + 1; r < n; r++) x[r] -= m[(size_t)c * n + r] * x[c];
    for (int c = n - 1; c >= 0; c--) {
      x[c] /= m[(size_t)c * n + c];
      for (int r = 0; r < c; r++) x[r] -= m[(size_t)c * n + r] * x[c];
    }
  }
  return true;
}

// out = alpha * a b (+ out when acc), column-major n x n
static void dense_mm(int n, const double* a, const double* b, double alpha, bool acc, double* out)
{
  for (int c = 0; c < n; c++) {
    double* o = out + (size_t)c * n;
    if (!acc) std::fill(o, o + n, 0.0);
    for (int k = 0; k < n; k++) {
      const double bk = alpha * b[(size_t)c * n + k];
      if (bk == 0.0) continue;
      const double* ak = a + (size_t)k * n;
      for (int r = 0; r < n; r++) o[r] += ak[r] * bk;
    }
  }
}

// Block structure of the coarse matrix: usable when every slot couples aggregates of one
// strip or of neighbouring strips and no strip holds more than CR_MAX aggregates.
static void cr_setup(Coarse& C, const std::vector<int>& agg_strip, int nstrips)
{
  C.cr_ok = C.cr_fac = false;
  if (getenv("RMX_COARSE_HOST")) return;
  std::vector<int> cnt(nstrips, 0);
  for (int v : agg_strip) cnt[v]++;
  int bs = 0;
  for (int v : cnt) bs = std::max(bs, v);
  if (bs < 1 || bs > CR_MAX) return;
  for (long long k = 0; k < C.nslot; k++)
    if (std::abs(agg_strip[C.slot_row[k]] - agg_strip[C.slot_col[k]]) > 1) return;
  const int nb = nstrips;
  C.nb = nb; C.bs = bs;
  C.blk_start.assign(nb + 1, 0);
  for (int i = 0; i < nb; i++) C.blk_start[i + 1] = C.blk_start[i] + cnt[i];
  for (int I = 0; I < C.nc; I++)                // aggregates are numbered strip by strip
    if (I < C.blk_start[agg_strip[I]] || I >= C.blk_start[agg_strip[I] + 1]) return;
  const long long b2 = (long long)bs * bs;
  C.slot_dst.resize(C.nslot);
  for (long long k = 0; k < C.nslot; k++) {
    const int I = C.slot_row[k], J = C.slot_col[k];
    const int bi = agg_strip[I], bj = agg_strip[J];
    const int r = I - C.blk_start[bi], c = J - C.blk_start[bj];
    const int which = bj - bi + 1;              // 0: B (i - 1), 1: A, 2: C (i + 1)
    C.slot_dst[k] = ((long long)which * nb + bi) * b2 + (long long)c * bs + r;
  }
  // matrix offsets of every level: eliminated blocks (A^-1, B, C), kept blocks (P, Q)
  C.lev_elim.clear(); C.lev_keep.clear();
  long long off = 0;
  int levels = 0;
  for (int s = 1; s < nb; s *= 2, levels++) {
    const int ne = (nb - s + 2 * s - 1) / (2 * s);
    const int nk = (nb + 2 * s - 1) / (2 * s);
    C.lev_elim.push_back(off); off += 3 * b2 * ne;
    C.lev_keep.push_back(off); off += 2 * b2 * nk;
  }
  C.levels = levels;
  C.top_off = off; off += b2;
  C.mats_n = off;
  C.h_mats.assign(off, 0.0);
  C.mats.alloc(off);
  C.w.alloc((size_t)nb * bs);
  C.d_blk_start.upload(C.blk_start, nullptr);
  C.d_slot_dst.upload(C.slot_dst, nullptr);
  HIPCHECK(hipDeviceSynchronize());
  C.cr_ok = true;
}

// Host factorisation of the reduction from the coarse entries cval (per slot).  Returns false
// when a reduced block is singular (the band LU then serves).
static bool cr_factor(Coarse& C, const double* cval)
{
  const int nb = C.nb, bs = C.bs;
  const long long b2 = (long long)bs * bs;
  std::vector<double> Bm(nb * b2, 0.0), Am(nb * b2, 0.0), Cm(nb * b2, 0.0);
  double* blocks[3] = {Bm.data(), Am.data(), Cm.data()};
  for (long long k = 0; k < C.nslot; k++) {
    const long long d = C.slot_dst[k];
    const int which = (int)(d / ((long long)nb * b2));
    blocks[which][d - (long long)which * nb * b2] = cval[k];
  }
  for (int i = 0; i < nb; i++) {                 // padding: identity
    const int m = C.blk_start[i + 1] - C.blk_start[i];
    for (int r = m; r < bs; r++) Am[i * b2 + (long long)r * bs + r] = 1.0;
  }
  std::vector<double> inv(nb * b2);
  std::atomic<bool> ok{true};
  int lev = 0;
  for (int s = 1; s < nb; s *= 2, lev++) {
    const int ne = (nb - s + 2 * s - 1) / (2 * s);
    const int nk = (nb + 2 * s - 1) / (2 * s);
    double* E = C.h_mats.data() + C.lev_elim[lev];
    double* K = C.h_mats.data() + C.lev_keep[lev];
    parallel_for(ne, [&](long long a, long long e) {
      for (long long k = a; k < e; k++) {
        const int j = (int)(s + 2 * s * k);
        double* Ai = E + 3 * b2 * k;
        if (!dense_inverse(bs, &Am[j * b2], Ai)) ok = false;
        std::copy(&Bm[j * b2], &Bm[j * b2] + b2, Ai + b2);
        std::copy(&Cm[j * b2], &Cm[j * b2] + b2, Ai + 2 * b2);
        std::copy(Ai, Ai + b2, &inv[j * b2]);
      }
    });
    if (!ok) return false;
    parallel_for(nk, [&](long long a, long long e) {
      std::vector<double> nB(b2), nC(b2);
      for (long long k = a; k < e; k++) {
        const int i = (int)(2 * s * k);
        double* P = K + 2 * b2 * k;
        double* Q = P + b2;
        std::fill(P, P + 2 * b2, 0.0);
        std::fill(nB.begin(), nB.end(), 0.0);
        std::fill(nC.begin(), nC.end(), 0.0);
        if (i - s >= 0) {
          dense_mm(bs, &Bm[i * b2], &inv[(i - s) * b2], 1.0, false, P);
          dense_mm(bs, P, &Cm[(i - s) * b2], -1.0, true, &Am[i * b2]);
          dense_mm(bs, P, &Bm[(i - s) * b2], -1.0, false, nB.data());
        }
        if (i + s < nb) {
          dense_mm(bs, &Cm[i * b2], &inv[(i + s) * b2], 1.0, false, Q);
          dense_mm(bs, Q, &Bm[(i + s) * b2], -1.0, true, &Am[i * b2]);
          dense_mm(bs, Q, &Cm[(i + s) * b2], -1.0, false, nC.data());
        }
        std::copy(nB.begin(), nB.end(), &Bm[i * b2]);
        std::copy(nC.begin(), nC.end(), &Cm[i * b2]);
      }
    });
  }
  return dense_inverse(bs, &Am[0], C.h_mats.data() + C.top_off);
}

// The factorisation on the device from the coarse entries on the device (C.cval): the same
// steps as cr_factor, a workgroup per block and level.  Returns false for a singular block.
static bool cr_factor_device(Ctx& c, Coarse& C)
{
  const int nb = C.nb, bs = C.bs;
  const long long b2 = (long long)bs * bs;
  hipStream_t st = c.stream;
  C.d_blocks.alloc(3 * nb * b2);
  C.d_inv.alloc(nb * b2);
  C.d_tmp.alloc(nb * b2 + 2 * b2);
  C.d_bad.alloc(1);
  HIPCHECK(hipMemsetAsync(C.d_blocks.p, 0, sizeof(double) * 3 * nb * b2, st));
  HIPCHECK(hipMemsetAsync(C.d_bad.p, 0, sizeof(int), st));
  double* Bm = C.d_blocks.p;
  double* Am = Bm + nb * b2;
  double* Cm = Am + nb * b2;
  k_cr_scatter<<<(int)((C.nslot + 255) / 256), 256, 0, st>>>(C.nslot, C.d_slot_dst.p, C.cval.p, Bm);
  k_cr_pad<<<nb, (bs + 63) / 64 * 64, 0, st>>>(nb, bs, C.d_blk_start.p, Am);
  const size_t lds = sizeof(double) * b2;
  static bool attr = false;
  if (!attr && lds > 65536) {
    HIPCHECK(hipFuncSetAttribute((const void*)k_cr_inv, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(double) * CR_MAX * CR_MAX)));
    attr = true;
  }
  int lev = 0;
  for (int s = 1; s < nb; s *= 2, lev++) {
    const int ne = (nb - s + 2 * s - 1) / (2 * s);
    const int nk = (nb + 2 * s - 1) / (2 * s);
    k_cr_inv<<<ne, 256, lds, st>>>(nb, bs, s, Am, Bm, Cm, C.mats.p + C.lev_elim[lev], C.d_inv.p,
                                   C.d_bad.p);
    k_cr_keep<<<nk, 256, 0, st>>>(nb, bs, s, Am, Bm, Cm, C.d_inv.p, C.mats.p + C.lev_keep[lev],
                                  C.d_tmp.p);
  }
  k_cr_inv<<<1, 256, lds, st>>>(nb, bs, 0, Am, Bm, Cm, C.mats.p + C.top_off, C.d_inv.p, C.d_bad.p);
  HIPCHECK(hipGetLastError());
  int bad = 0;
  HIPCHECK(hipMemcpyAsync(&bad, C.d_bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return bad == 0;
}

static void cr_apply(Ctx& c, Coarse& C)
{
  const int nb = C.nb, bs = C.bs;
  const int nt = (bs + 63) / 64 * 64;           // a thread per block row, whole waves
  hipStream_t st = c.stream;
  k_cr_in<<<nb, nt, 0, st>>>(nb, bs, C.d_blk_start.p, C.rc.p, C.w.p);
  int lev = 0;
  for (int s = 1; s < nb; s *= 2, lev++) {
    const int nk = (nb + 2 * s - 1) / (2 * s);
    k_cr_fwd<<<nk, nt, 0, st>>>(nb, bs, s, C.mats.p + C.lev_keep[lev], C.w.p);
  }
  k_cr_bwd<<<1, nt, 0, st>>>(nb, bs, 0, C.mats.p + C.top_off, C.w.p);
  for (int l = C.levels - 1; l >= 0; l--) {
    const int s = 1 << l;
    const int ne = (nb - s + 2 * s - 1) / (2 * s);
    k_cr_bwd<<<ne, nt, 0, st>>>(nb, bs, s, C.mats.p + C.lev_elim[l], C.w.p);
  }
  k_cr_out<<<nb, nt, 0, st>>>(nb, bs, C.d_blk_start.p, C.w.p, C.zc.p);
  HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------

// Coarse segments per strip: the chosen count, or automatically one per 256 rows of a strip
// (aggregates of about 85 rows per variable), at most 24.  Measured with FGMRES(30) to an
// FP64 residual of 1e-8: the 9.2 M-equation meander (strips of 18,000 rows) 874 iterations
// with 16 segments, 707 with 24; a 13,118-equation channel (strips of 2,186 rows) 267 with
// 8, 1,144 with 16, 3,225 with 24; a 231,442-equation meander reach (strips of 2,048 rows)
// 0.84 s per Newton step with 8, 1.06 s with 24.
int coarse_segments_of(const Ctx& c)
{
  if (c.coarse_segments >= 0) return c.coarse_segments;
  const long long rows = c.n / std::max(1, c.silu.nblk);
  return (int)std::max<long long>(1, std::min<long long>(24, rows / 256));
}

void coarse_setup(Ctx& c, const std::vector<int>& rp, const std::vector<int>& cl,
                  const std::vector<int>& group, int nseg)
{
  Coarse& C = c.coarse;
  C.ready = C.fac = false;
  const int n = c.n;
  const Ilu& I = c.silu;
  const std::vector<int> seg = nseg > 1 ? cut_segments(I.bstart, group, c.hm, nseg)
                                        : std::vector<int>(n, 0);
  // variable of every row from the numbering (eqno[3*node + j], j = U, V, H)
  std::vector<int> var(n, 0);
  for (size_t k = 0; k < c.hm.eqno.size(); k++)
    if (c.hm.eqno[k] >= 0) var[c.hm.eqno[k]] = (int)(k % 3);
  // aggregates, numbered in (strip, segment, variable) order, empty ones dropped
  const long long nraw = (long long)I.nblk * nseg * 3;
  std::vector<int> raw(n), idx(nraw, -1);
  for (int b = 0; b < I.nblk; b++)
    for (int i = I.bstart[b]; i < I.bstart[b + 1]; i++) {
      raw[i] = (int)(((long long)b * nseg + seg[i]) * 3 + var[i]);
      idx[raw[i]] = 0;
    }
  int nc = 0;
  std::vector<int> agg_strip;
  for (long long k = 0; k < nraw; k++)
    if (idx[k] == 0) {
      idx[k] = nc++;
      agg_strip.push_back((int)(k / (3LL * nseg)));
    }
  std::vector<int> agg(n);
  for (int i = 0; i < n; i++) agg[i] = idx[raw[i]];
  // band of the coarse matrix, then coarse entries (I, J) in row-major order via a map
  // indexed by (I, J - I + kl)
  int kl = 0, ku = 0;
  for (int i = 0; i < n; i++)
    for (int q = rp[i]; q < rp[i + 1]; q++) {
      const int d = agg[cl[q]] - agg[i];
      kl = std::max(kl, -d); ku = std::max(ku, d);
    }
  const long long bw = (long long)kl + ku + 1;
  if ((long long)nc * bw > (1LL << 28))
    fail(RMX_ERR_RANGE, "coarse space too large (" + std::to_string(nc) + " aggregates, band " +
                        std::to_string(bw) + ")");
  std::vector<int> key_slot((size_t)nc * bw, -1);
  auto key = [&](int I, int J) { return (size_t)I * bw + (size_t)(J - I + kl); };
  for (int i = 0; i < n; i++)
    for (int q = rp[i]; q < rp[i + 1]; q++) key_slot[key(agg[i], agg[cl[q]])] = 0;
  C.slot_row.clear(); C.slot_col.clear();
  for (size_t k = 0; k < key_slot.size(); k++)
    if (key_slot[k] == 0) {
      key_slot[k] = (int)C.slot_row.size();
      const int r = (int)(k / bw), col = r + (int)(k % bw) - kl;
      C.slot_row.push_back(r); C.slot_col.push_back(col);
    }
  const long long nslot = (long long)C.slot_row.size();
  const long long nnz = rp[n];
  std::vector<long long> sptr(nslot + 1, 0);
  std::vector<int> slot_of(nnz);
  parallel_for(n, [&](long long a, long long e) {
    for (long long i = a; i < e; i++)
      for (int q = rp[i]; q < rp[i + 1]; q++) slot_of[q] = key_slot[key(agg[i], agg[cl[q]])];
  });
  for (long long q = 0; q < nnz; q++) sptr[slot_of[q] + 1]++;
  for (long long s = 0; s < nslot; s++) sptr[s + 1] += sptr[s];
  std::vector<int> perm(nnz);
  {
    std::vector<long long> fill(sptr.begin(), sptr.end() - 1);
    for (long long q = 0; q < nnz; q++) perm[fill[slot_of[q]]++] = (int)q;
  }
  // rows of every aggregate, ascending
  std::vector<int> aptr(nc + 1, 0), arows(n);
  for (int i = 0; i < n; i++) aptr[agg[i] + 1]++;
  for (int k = 0; k < nc; k++) aptr[k + 1] += aptr[k];
  {
    std::vector<int> fill(aptr.begin(), aptr.end() - 1);
    for (int i = 0; i < n; i++) arows[fill[agg[i]]++] = i;
  }
  // A P structure: per row its distinct column aggregates (ascending) and the row's entries
  // grouped per aggregate
  std::vector<int> apc(n + 1, 0);
  parallel_for(n, [&](long long a, long long e) {
    std::vector<int> tmp;
    for (long long i = a; i < e; i++) {
      tmp.clear();
      for (int q = rp[i]; q < rp[i + 1]; q++) tmp.push_back(agg[cl[q]]);
      std::sort(tmp.begin(), tmp.end());
      apc[i + 1] = (int)(std::unique(tmp.begin(), tmp.end()) - tmp.begin());
    }
  });
  for (int i = 0; i < n; i++) apc[i + 1] += apc[i];
  const long long nap = apc[n];
  std::vector<int> ap_agg(nap), src_ptr(nap + 1), src(rp[n]);
  parallel_for(n, [&](long long a, long long e) {
    std::vector<std::pair<int, int>> tmp;
    for (long long i = a; i < e; i++) {
      tmp.clear();
      for (int q = rp[i]; q < rp[i + 1]; q++) tmp.emplace_back(agg[cl[q]], q);
      std::stable_sort(tmp.begin(), tmp.end(),
                       [](const std::pair<int, int>& x, const std::pair<int, int>& y) { return x.first < y.first; });
      long long k = apc[i] - 1;
      int pos = rp[i];
      for (size_t t = 0; t < tmp.size(); t++) {
        if (t == 0 || tmp[t].first != tmp[t - 1].first) { k++; ap_agg[k] = tmp[t].first; src_ptr[k] = pos; }
        src[pos++] = tmp[t].second;
      }
    }
  });
  src_ptr[nap] = rp[n];
  C.nap = nap;
  C.ap_ptr.upload(apc, c.stream); C.ap_agg.upload(ap_agg, c.stream);
  C.ap_src_ptr.upload(src_ptr, c.stream); C.ap_src.upload(src, c.stream);
  C.apv.alloc(nap);
  C.nc = nc; C.kl = kl; C.ku = ku; C.nslot = nslot;
  hipStream_t s = c.stream;
  C.agg.upload(agg, s); C.agg_ptr.upload(aptr, s); C.agg_rows.upload(arows, s);
  C.slot_ptr.upload(sptr, s); C.perm.upload(perm, s);
  C.cval.alloc(nslot); C.rc.alloc(nc); C.zc.alloc(nc); C.t.alloc(n);
  if (C.h_buf) { (void)hipHostFree(C.h_buf); C.h_buf = nullptr; }
  HIPCHECK(hipHostMalloc(&C.h_buf, sizeof(double) * std::max<long long>(nc, nslot),
                         hipHostMallocDefault));
  HIPCHECK(hipStreamSynchronize(s));
  cr_setup(C, agg_strip, I.nblk);
  C.ready = true;
  if (getenv("RMX_DEBUG"))
    fprintf(stderr, "[rmx] coarse space: %d aggregates (%d strips x %d segments x U/V/H), "
            "%lld entries, band kl %d ku %d; %s\n", nc, I.nblk, nseg, nslot, kl, ku,
            C.cr_ok ? "device block cyclic reduction" : "host band LU");
}

void coarse_factor(Ctx& c, int prec)
{
  Coarse& C = c.coarse;
  if (!C.ready) fail(RMX_ERR_ARG, "coarse space: no schedule (set up from a mesh)");
  const int blocks = (int)((C.nslot * 64 + 255) / 256);
  const int gap = (int)((C.nap + 255) / 256);
  if (prec == RMX_F32) {
    if (!c.have32) fail(RMX_ERR_ARG, "coarse space: no FP32 matrix");
    k_coarse_vals<float><<<blocks, 256, 0, c.stream>>>(C.nslot, C.slot_ptr.p, C.perm.p, c.A32.p, C.cval.p);
    k_ap_vals<float><<<gap, 256, 0, c.stream>>>(C.nap, C.ap_src_ptr.p, C.ap_src.p, c.A32.p, C.apv.p);
  } else {
    if (!c.have64) fail(RMX_ERR_ARG, "coarse space: no FP64 matrix");
    k_coarse_vals<double><<<blocks, 256, 0, c.stream>>>(C.nslot, C.slot_ptr.p, C.perm.p, c.A64.p, C.cval.p);
    k_ap_vals<double><<<gap, 256, 0, c.stream>>>(C.nap, C.ap_src_ptr.p, C.ap_src.p, c.A64.p, C.apv.p);
  }
  HIPCHECK(hipGetLastError());
  C.fac = false;
  C.cr_fac = false;
  // the block cyclic reduction factored on the device (RMX_COARSE_HOST_FACTOR: on the host)
  if (C.cr_ok && !getenv("RMX_COARSE_HOST_FACTOR") && cr_factor_device(c, C)) {
    C.cr_fac = true;
    C.fac = true;
    C.fac_prec = prec;
    return;
  }
  HIPCHECK(hipMemcpyAsync(C.h_buf, C.cval.p, sizeof(double) * C.nslot, hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  if (C.cr_ok && cr_factor(C, C.h_buf)) {
    HIPCHECK(hipMemcpyAsync(C.mats.p, C.h_mats.data(), sizeof(double) * C.mats_n,
                            hipMemcpyHostToDevice, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    C.cr_fac = true;
    C.fac = true;
    C.fac_prec = prec;
    return;
  }
  const int ldab = 2 * C.kl + C.ku + 1, kv = C.kl + C.ku;
  C.band.assign((size_t)ldab * C.nc, 0.0);
  for (long long k = 0; k < C.nslot; k++)
    bel(C.band, ldab, kv, C.slot_row[k], C.slot_col[k]) = C.h_buf[k];
  C.fac = false;
  if (!band_lu(C.nc, C.kl, C.ku, C.band, C.ipiv))
    fail(RMX_ERR_PIVOT, "coarse space: singular coarse matrix");
  C.fac = true;
  C.fac_prec = prec;
}

void coarse_apply(Ctx& c, int prec, const double* r, double* z)
{
  Coarse& C = c.coarse;
  if (!C.fac || C.fac_prec != prec) fail(RMX_ERR_ARG, "coarse space not factored in this precision");
  const int n = c.n, g = (n + 255) / 256;
  // bytes: restriction (row lists, r), coarse solve (its matrices), A P residual (A P
  // entries and aggregates, r, t)
  const int kt = kt_begin(c, KT_COARSE, 4.0 * (C.nc + 1) + 12.0 * n + 8.0 * C.nc +
                                        (C.cr_fac ? 8.0 * (double)C.mats_n : 0.0) +
                                        12.0 * (double)C.nap + 4.0 * (n + 1) + 16.0 * n);
  k_restrict<<<(C.nc * 64 + 255) / 256, 256, 0, c.stream>>>(C.nc, C.agg_ptr.p, C.agg_rows.p, r, C.rc.p);
  if (C.cr_fac) {
    cr_apply(c, C);
  } else {
    HIPCHECK(hipMemcpyAsync(C.h_buf, C.rc.p, sizeof(double) * C.nc, hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    band_solve(C.nc, C.kl, C.ku, C.band, C.ipiv, C.h_buf);
    HIPCHECK(hipMemcpyAsync(C.zc.p, C.h_buf, sizeof(double) * C.nc, hipMemcpyHostToDevice, c.stream));
  }
  k_ap_resid<<<g, 256, 0, c.stream>>>(n, C.ap_ptr.p, C.ap_agg.p, C.apv.p, C.zc.p, r, C.t.p);
  kt_end(c, kt);
  if (!ilu_apply(c, c.silu, prec, C.t.p, z, nullptr, C.agg.p, C.zc.p))
    k_add_prolong<<<g, 256, 0, c.stream>>>(n, C.agg.p, C.zc.p, z);
  HIPCHECK(hipGetLastError());
}

}  // namespace rmx
