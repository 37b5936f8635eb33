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

void coarse_setup(Ctx& c, const std::vector<int>& rp, const std::vector<int>& cl,
                  const std::vector<int>& group)
{
  Coarse& C = c.coarse;
  C.ready = C.fac = false;
  const int n = c.n;
  const Ilu& I = c.silu;
  const int nseg = c.coarse_segments;
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
  for (long long k = 0; k < nraw; k++)
    if (idx[k] == 0) idx[k] = nc++;
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
  C.ready = true;
  if (getenv("RMX_DEBUG"))
    fprintf(stderr, "[rmx] coarse space: %d aggregates (%d strips x %d segments x U/V/H), "
            "%lld entries, band kl %d ku %d\n", nc, I.nblk, nseg, nslot, kl, ku);
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
  HIPCHECK(hipMemcpyAsync(C.h_buf, C.cval.p, sizeof(double) * C.nslot, hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
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
  k_restrict<<<(C.nc * 64 + 255) / 256, 256, 0, c.stream>>>(C.nc, C.agg_ptr.p, C.agg_rows.p, r, C.rc.p);
  HIPCHECK(hipMemcpyAsync(C.h_buf, C.rc.p, sizeof(double) * C.nc, hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  band_solve(C.nc, C.kl, C.ku, C.band, C.ipiv, C.h_buf);
  HIPCHECK(hipMemcpyAsync(C.zc.p, C.h_buf, sizeof(double) * C.nc, hipMemcpyHostToDevice, c.stream));
  k_ap_resid<<<g, 256, 0, c.stream>>>(n, C.ap_ptr.p, C.ap_agg.p, C.apv.p, C.zc.p, r, C.t.p);
  if (!ilu_apply(c, c.silu, prec, C.t.p, z, nullptr, C.agg.p, C.zc.p))
    k_add_prolong<<<g, 256, 0, c.stream>>>(n, C.agg.p, C.zc.p, z);
  HIPCHECK(hipGetLastError());
}

}  // namespace rmx
