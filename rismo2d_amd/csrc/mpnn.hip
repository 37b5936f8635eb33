// Learned preconditioner: a message-passing network on the mesh graph, evaluated and
// trained on the device with this file's kernels.  No reference counterpart.
//
// Graph: mesh nodes; an edge joins two nodes of one element (both directions).  The
// network sees the state of the Newton system, not the residual, so the preconditioner it
// defines is linear in the residual and FGMRES keeps its Krylov structure.  Node features
// (8): u, v, depth/4, bed/4, 100 cf, speed, open-boundary flag, rotated-node flag.
//
//     h0 = relu(f W_in + b_in)                  (FP32, stored BF16)
//     6 x:  a = mean of h over the neighbours   (FP32 sums of BF16, stored BF16)
//           p = [h, a] W_l + b_l                (BF16 MFMA 16x16x32, FP32 accumulate)
//           h = h + relu(p)                     (stored BF16)
//     out = h W_out + b_out                     (FP32, 18 per node): two 3x3 blocks D_i, E_i
//
// The correction of a residual r: q = r / max_j |a_ij| (row-scaled), q_i its (u, v, h)
// components at node i, qbar_i their mean over the neighbours, and
//     n(r)_i = D_i q_i + E_i qbar_i.
// As a preconditioner (RMX_PREC_MPNN) it corrects what the two-level preconditioner M2
// leaves: z2 = M2^-1 r, r2 = r - A z2, z = z2 + n(r2).  The network runs once per Newton
// system; at the first application of a system the correction is tested on that vector
// and dropped for the whole system if it does not lower the residual.  Training
// minimises the preconditioned residual ||r - A n(r)||^2 / ||r||^2 over stored samples
// (r and the matrix A of the Newton system it came from); its gradient reaches D, E
// through A^T.  Gradients by the backward kernels below (weight gradients as per-block
// partial sums reduced in a fixed order, so they are deterministic), Adam on the FP32
// master weights, BF16 copies refreshed after each step.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "rmx_internal.h"

namespace rmx {

typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// gfx950's BF16 MFMA with K = 32 per instruction (16 x 16 output tile, FP32 accumulation):
// lane l holds 8 of the K values of row (A) or column (B) l % 16, the same 8 in both
__device__ __forceinline__ f4 mfma_k32(s8 a, s8 b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a),
                                                 __builtin_bit_cast(b8, b), c, 0, 0, 0);
}

constexpr int MC = Mpnn::C, ML = Mpnn::L, MF = Mpnn::F, MO = Mpnn::O;
constexpr int OFF_WIN = 0, OFF_BIN = MF * MC;
constexpr int LAYER_SZ = 2 * MC * MC + MC;
constexpr int OFF_L0 = OFF_BIN + MC;
constexpr int OFF_WOUT = OFF_L0 + ML * LAYER_SZ, OFF_BOUT = OFF_WOUT + MC * MO;
constexpr int NPARAM = OFF_BOUT + MO;
static_assert(NPARAM == 51282, "parameter layout");
constexpr int CHUNK = 64;          // nodes per block of the small reductions
constexpr int WCHUNK = 256;        // nodes per block of the layer weight gradients

__device__ __forceinline__ uint16_t f2bf(float f)
{
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---------------------------------------------------------------------------------------
// features, targets, outputs

__global__ void k_feat(int np, const double* __restrict__ U, const double* __restrict__ V,
                       const double* __restrict__ S, const double* __restrict__ z,
                       const double* __restrict__ cf, const uint32_t* __restrict__ flag,
                       float* __restrict__ feat)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  float* f = feat + (size_t)i * MF;
  f[0] = (float)U[i];
  f[1] = (float)V[i];
  f[2] = (float)(fmax(S[i] - z[i], 0.0) * 0.25);
  f[3] = (float)(z[i] * 0.25);
  f[4] = (float)(cf[i] * 100.0);
  f[5] = (float)sqrt(U[i] * U[i] + V[i] * V[i]);
  f[6] = (flag[i] & (dInlet | dOutlet)) ? 1.0f : 0.0f;
  f[7] = (flag[i] & dRotat) ? 1.0f : 0.0f;
}

// q components at node i and their neighbour mean
__device__ __forceinline__ void node_q(int i, const int* __restrict__ eqno, const int* __restrict__ ptr,
                                       const int* __restrict__ adj, const double* __restrict__ q,
                                       double qi[3], double qb[3])
{
  for (int w = 0; w < 3; w++) {
    const int e = eqno[3 * i + w];
    qi[w] = e >= 0 ? q[e] : 0.0;
    qb[w] = 0.0;
  }
  const int a = ptr[i], b = ptr[i + 1];
  for (int k = a; k < b; k++) {
    const int j = adj[k];
    for (int w = 0; w < 3; w++) {
      const int e = eqno[3 * j + w];
      if (e >= 0) qb[w] += q[e];
    }
  }
  if (b > a)
    for (int w = 0; w < 3; w++) qb[w] /= (double)(b - a);
}

// n_i = D_i q_i + E_i qbar_i (out: [np][18], D then E, row-major 3x3)
__global__ void k_lin_apply(int np, const int* __restrict__ eqno, const int* __restrict__ ptr,
                            const int* __restrict__ adj, const float* __restrict__ out,
                            const double* __restrict__ q, double* __restrict__ nv)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double qi[3], qb[3];
  node_q(i, eqno, ptr, adj, q, qi, qb);
  const float* o = out + (size_t)i * MO;
  for (int v = 0; v < 3; v++) {
    const int e = eqno[3 * i + v];
    if (e < 0) continue;
    double s = 0.0;
    for (int w = 0; w < 3; w++) s += (double)o[3 * v + w] * qi[w] + (double)o[9 + 3 * v + w] * qb[w];
    nv[e] = s;
  }
}

// gout[i] = d loss / d (D_i, E_i) from g = d loss / d n (equation vector); with `add` the
// gradient is added to what gout holds
__global__ void k_lin_grad(int np, const int* __restrict__ eqno, const int* __restrict__ ptr,
                           const int* __restrict__ adj, const double* __restrict__ q,
                           const double* __restrict__ g, int add, float* __restrict__ gout)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double qi[3], qb[3];
  node_q(i, eqno, ptr, adj, q, qi, qb);
  float* o = gout + (size_t)i * MO;
  for (int v = 0; v < 3; v++) {
    const int e = eqno[3 * i + v];
    const double gv = e >= 0 ? g[e] : 0.0;
    for (int w = 0; w < 3; w++) {
      const float a = (float)(gv * qi[w]), b = (float)(gv * qb[w]);
      o[3 * v + w] = add ? o[3 * v + w] + a : a;
      o[9 + 3 * v + w] = add ? o[9 + 3 * v + w] + b : b;
    }
  }
}

// res = b - A x for a stored sample's matrix values (FP64, 8 lanes per row)
__global__ void __launch_bounds__(256) k_resid_s(int n, const int* __restrict__ rowptr,
                                                 const int* __restrict__ cols,
                                                 const double* __restrict__ A,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ b,
                                                 double* __restrict__ res)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t >> 3), lane = (int)(t & 7);
  if (row >= n) return;
  double s = 0.0;
  for (int p = rowptr[row] + lane; p < rowptr[row + 1]; p += 8) s += A[p] * x[cols[p]];
  for (int o = 4; o > 0; o >>= 1) s += __shfl_xor(s, o, 8);
  if (lane == 0) res[row] = b[row] - s;
}

// y = scale * A^T x with A given in CSR order and its transpose as column pointers, row
// indices and the CSR position of every entry
__global__ void __launch_bounds__(256) k_spmv_t(int n, const int* __restrict__ tptr,
                                                const int* __restrict__ trow,
                                                const int* __restrict__ tperm,
                                                const double* __restrict__ A,
                                                const double* __restrict__ x, double scale,
                                                double* __restrict__ y)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = (int)(t >> 3), lane = (int)(t & 7);
  if (j >= n) return;
  double s = 0.0;
  for (int k = tptr[j] + lane; k < tptr[j + 1]; k += 8) s += A[tperm[k]] * x[trow[k]];
  for (int o = 4; o > 0; o >>= 1) s += __shfl_xor(s, o, 8);
  if (lane == 0) y[j] = scale * s;
}

// ---------------------------------------------------------------------------------------
// forward

__global__ void k_in(int np, const float* __restrict__ feat, const float* __restrict__ th,
                     uint16_t* __restrict__ H0, float* __restrict__ Q)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(t / MC), c = (int)(t % MC);
  if (i >= np) return;
  float q = th[OFF_BIN + c];
  for (int f = 0; f < MF; f++) q += feat[(size_t)i * MF + f] * th[OFF_WIN + f * MC + c];
  if (Q) Q[(size_t)i * MC + c] = q;
  H0[(size_t)i * MC + c] = f2bf(fmaxf(q, 0.0f));
}

// a_i = mean over neighbours of h (one wave per node, lane = channel)
__global__ void __launch_bounds__(256) k_agg(int np, const int* __restrict__ ptr,
                                             const int* __restrict__ adj,
                                             const uint16_t* __restrict__ H,
                                             uint16_t* __restrict__ A)
{
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int c = threadIdx.x & 63;
  if (i >= np) return;
  const int a = ptr[i], b = ptr[i + 1];
  float s = 0.0f;
  for (int k = a; k < b; k++) s += bf2f(H[(size_t)adj[k] * MC + c]);
  A[(size_t)i * MC + c] = f2bf(b > a ? s / (float)(b - a) : 0.0f);
}

// h_out = h + relu([h, a] W + b): 64 nodes per block, 16 per wave, 4 output tiles of 16
// channels per wave, K = 128 in 4 MFMA steps.  Wt = W^T (BF16, [64][128]).
__global__ void __launch_bounds__(256) k_layer_fwd(int np, const uint16_t* __restrict__ H,
                                                   const uint16_t* __restrict__ A,
                                                   const uint16_t* __restrict__ Wt,
                                                   const float* __restrict__ bias,
                                                   uint16_t* __restrict__ Hout,
                                                   float* __restrict__ P)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nb = blockIdx.x * 64 + wave * 16;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int node = nb + r16;
  f4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    const int k0 = ks * 32 + kq;
    s8 a = s8{0, 0, 0, 0, 0, 0, 0, 0};
    if (node < np)
      a = *reinterpret_cast<const s8*>(k0 < MC ? H + (size_t)node * MC + k0
                                               : A + (size_t)node * MC + (k0 - MC));
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const s8 b = *reinterpret_cast<const s8*>(Wt + (size_t)(t * 16 + r16) * (2 * MC) + k0);
      acc[t] = mfma_k32(a, b, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int nd = nb + (lane >> 4) * 4 + j;
      const int col = t * 16 + r16;
      if (nd < np) {
        const float p = acc[t][j] + bias[col];
        const size_t o = (size_t)nd * MC + col;
        Hout[o] = f2bf(bf2f(H[o]) + fmaxf(p, 0.0f));
        if (P) P[o] = p;
      }
    }
}

// out = h W_out + b_out (FP32), one thread per node
__global__ void k_out(int np, const uint16_t* __restrict__ H, const float* __restrict__ th,
                      float* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  float o[MO];
  for (int v = 0; v < MO; v++) o[v] = th[OFF_BOUT + v];
  for (int c = 0; c < MC; c++) {
    const float h = bf2f(H[(size_t)i * MC + c]);
    for (int v = 0; v < MO; v++) o[v] += h * th[OFF_WOUT + c * MO + v];
  }
  for (int v = 0; v < MO; v++) out[(size_t)i * MO + v] = o[v];
}

// ---------------------------------------------------------------------------------------
// backward

// gh = gout W_out^T
__global__ void k_gh_out(int np, const float* __restrict__ gout, const float* __restrict__ th,
                         float* __restrict__ gh)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(t / MC), c = (int)(t % MC);
  if (i >= np) return;
  float s = 0.0f;
  for (int v = 0; v < MO; v++) s += gout[(size_t)i * MO + v] * th[OFF_WOUT + c * MO + v];
  gh[(size_t)i * MC + c] = s;
}

// gp = bf16(gh * (p > 0))
__global__ void k_gp(long long m, const float* __restrict__ gh, const float* __restrict__ P,
                     uint16_t* __restrict__ gp)
{
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < m;
       t += (long long)gridDim.x * blockDim.x)
    gp[t] = f2bf(P[t] > 0.0f ? gh[t] : 0.0f);
}

// gq = gh * (q > 0), in place
__global__ void k_gq(long long m, float* __restrict__ gh, const float* __restrict__ Q)
{
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < m;
       t += (long long)gridDim.x * blockDim.x)
    if (!(Q[t] > 0.0f)) gh[t] = 0.0f;
}

// gx = gp W^T ([np x 64] x [64 x 128]): gh += gx[:, :64] (in place), ga = gx[:, 64:].
// Wb = W (BF16, [128][64]).
__global__ void __launch_bounds__(256) k_layer_bwd(int np, const uint16_t* __restrict__ gp,
                                                   const uint16_t* __restrict__ Wb,
                                                   float* __restrict__ gh,
                                                   float* __restrict__ ga)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nb = blockIdx.x * 64 + wave * 16;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int node = nb + r16;
  f4 acc[8];
#pragma unroll
  for (int t = 0; t < 8; t++) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    const int k0 = ks * 32 + kq;
    s8 a = s8{0, 0, 0, 0, 0, 0, 0, 0};
    if (node < np) a = *reinterpret_cast<const s8*>(gp + (size_t)node * MC + k0);
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const s8 b = *reinterpret_cast<const s8*>(Wb + (size_t)(t * 16 + r16) * MC + k0);
      acc[t] = mfma_k32(a, b, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < 8; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int nd = nb + (lane >> 4) * 4 + j;
      const int col = t * 16 + r16;
      if (nd < np) {
        if (col < MC) gh[(size_t)nd * MC + col] += acc[t][j];
        else ga[(size_t)nd * MC + (col - MC)] = acc[t][j];
      }
    }
}

// gh_j += sum over neighbours i of ga_i / deg_i (the transpose of the mean aggregation;
// the graph is symmetric)
__global__ void __launch_bounds__(256) k_aggT(int np, const int* __restrict__ ptr,
                                              const int* __restrict__ adj,
                                              const float* __restrict__ ga,
                                              float* __restrict__ gh)
{
  const int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int c = threadIdx.x & 63;
  if (j >= np) return;
  float s = 0.0f;
  for (int k = ptr[j]; k < ptr[j + 1]; k++) {
    const int i = adj[k];
    const int d = ptr[i + 1] - ptr[i];
    s += ga[(size_t)i * MC + c] / (float)d;
  }
  gh[(size_t)j * MC + c] += s;
}

// partial[blk][m][n] = sum over the block's nodes of [h, a][node][m] * gp[node][n]
// (MFMA, K = nodes): 32 output tiles, 8 per wave.
__global__ void __launch_bounds__(256) k_wgrad(int np, const uint16_t* __restrict__ H,
                                               const uint16_t* __restrict__ A,
                                               const uint16_t* __restrict__ gp,
                                               float* __restrict__ partial)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int n0 = blockIdx.x * WCHUNK;
  f4 acc[2][4];
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int t = 0; t < 4; t++) acc[u][t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int kk = 0; kk < WCHUNK; kk += 32) {
    s8 a[2], b[4];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int m = (wave * 2 + u) * 16 + r16;
      const uint16_t* X = m < MC ? H : A;
      const int mc = m < MC ? m : m - MC;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int nd = n0 + kk + kq + q;
        a[u][q] = nd < np ? (short)X[(size_t)nd * MC + mc] : (short)0;
      }
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int n = t * 16 + r16;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int nd = n0 + kk + kq + q;
        b[t][q] = nd < np ? (short)gp[(size_t)nd * MC + n] : (short)0;
      }
    }
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int t = 0; t < 4; t++)
        acc[u][t] = mfma_k32(a[u], b[t], acc[u][t]);
  }
  float* out = partial + (size_t)blockIdx.x * (2 * MC * MC);
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int m = (wave * 2 + u) * 16 + (lane >> 4) * 4 + j;
        const int n = t * 16 + r16;
        out[m * MC + n] = acc[u][t][j];
      }
}

// Small outer-product reductions (input / output layers, biases):
// partial[blk][m][o] = sum over the block's nodes of x(node, m) * g(node, o), with
// x(node, mx) = 1 (the bias row).  XT / GT: float or BF16 (uint16_t).
template <class T> __device__ __forceinline__ float ld(const T* p, size_t i);
template <> __device__ __forceinline__ float ld<float>(const float* p, size_t i) { return p[i]; }
template <> __device__ __forceinline__ float ld<uint16_t>(const uint16_t* p, size_t i) { return bf2f(p[i]); }

template <class XT, class GT>
__global__ void __launch_bounds__(256) k_outer(int np, const XT* __restrict__ X, int mx,
                                               const GT* __restrict__ G, int ng,
                                               float* __restrict__ partial)
{
  // tp threads share a pair, each summing every tp-th node of the block's CHUNK; their sums
  // are added in a fixed order
  __shared__ float s_part[256];
  const int pairs = (mx + 1) * ng;
  int tp = 1;
  while (tp * 2 * pairs <= 256) tp *= 2;
  const int t = threadIdx.x, sub = t % tp;
  const int n0 = blockIdx.x * CHUNK, n1 = min(np, n0 + CHUNK);
  for (int base = 0; base < pairs; base += 256 / tp) {
    const int pr = base + t / tp;
    float s = 0.0f;
    if (pr < pairs) {
      const int m = pr / ng, o = pr % ng;
      if (m < mx)
        for (int i = n0 + sub; i < n1; i += tp)
          s += ld<XT>(X, (size_t)i * mx + m) * ld<GT>(G, (size_t)i * ng + o);
      else
        for (int i = n0 + sub; i < n1; i += tp) s += ld<GT>(G, (size_t)i * ng + o);
    }
    s_part[t] = s;
    __syncthreads();
    if (sub == 0 && pr < pairs) {
      float a = 0.0f;
      for (int q = 0; q < tp; q++) a += s_part[t + q];
      partial[(size_t)blockIdx.x * pairs + pr] = a;
    }
    __syncthreads();
  }
}

// dst[map(p)] = sum over blocks of partial[b][p] (fixed order); map: p = m*ng + o with
// m < mx -> offW + m*ng + o, m == mx -> offB + o
__global__ void k_reduce(int nblk, int pairs, const float* __restrict__ partial, int mx, int ng,
                         int offW, int offB, float* __restrict__ grad)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pairs) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; b++) s += partial[(size_t)b * pairs + p];
  const int m = p / ng, o = p % ng;
  grad[m < mx ? offW + m * ng + o : offB + o] = s;
}

// first stage of the same sum for many blocks: part2[y][p] = sum of partial[b][p] over the
// `per` consecutive blocks of slice y (four waves over interleaved blocks, combined in LDS in
// a fixed order); 64 entries per workgroup
__global__ void __launch_bounds__(256) k_reduce_slices(int nblk, int pairs, int per,
                                                       const float* __restrict__ partial,
                                                       float* __restrict__ part2)
{
  __shared__ float sh[4][64];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + l;
  const int b0 = blockIdx.y * per, b1 = min(nblk, b0 + per);
  float s = 0.0f;
  if (p < pairs)
    for (int b = b0 + w; b < b1; b += 4) s += partial[(size_t)b * pairs + p];
  sh[w][l] = s;
  __syncthreads();
  if (w == 0 && p < pairs) part2[(size_t)blockIdx.y * pairs + p] = ((sh[0][l] + sh[1][l]) + sh[2][l]) + sh[3][l];
}

__global__ void k_adam(int n, float* __restrict__ th, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                       float c1, float c2, float eps)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  m[i] = b1 * m[i] + (1.0f - b1) * gi;
  v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
  th[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
}

// BF16 copies of the layer weights: wb = W ([128][64]), wt = W^T ([64][128])
__global__ void k_refresh(const float* __restrict__ th, uint16_t* __restrict__ wt,
                          uint16_t* __restrict__ wb)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ML * 2 * MC * MC) return;
  const int l = t / (2 * MC * MC), r = t % (2 * MC * MC);
  const int m = r / MC, c = r % MC;
  const uint16_t w = f2bf(th[OFF_L0 + l * LAYER_SZ + m * MC + c]);
  wb[(size_t)l * 2 * MC * MC + m * MC + c] = w;
  wt[(size_t)l * 2 * MC * MC + c * 2 * MC + m] = w;
}

__global__ void k_add_v(int n, const double* __restrict__ nv, double* __restrict__ z)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] += nv[i];
}

// ---------------------------------------------------------------------------------------
// host side

static int gb(long long n, int b = 256) { return (int)((n + b - 1) / b); }

// grad[map(p)] = sum over nblk blocks of partial[b][p], in a fixed order: slices of up to
// 48 blocks summed first (k_reduce_slices), then the slices (k_reduce)
static void reduce_blocks(Ctx& c, int nblk, int pairs, const float* partial, int mx, int ng,
                          int offW, int offB, float* grad)
{
  constexpr int PER = 48;
  const int ns = gb(nblk, PER);
  if (ns <= 1) {
    k_reduce<<<gb(pairs), 256, 0, c.stream>>>(nblk, pairs, partial, mx, ng, offW, offB, grad);
    return;
  }
  DBuf<float>& P2 = c.mpnn.part2;
  if (P2.n < (size_t)ns * pairs) P2.alloc((size_t)ns * pairs);
  k_reduce_slices<<<dim3(gb(pairs, 64), ns), 256, 0, c.stream>>>(nblk, pairs, PER, partial, P2.p);
  k_reduce<<<gb(pairs), 256, 0, c.stream>>>(ns, pairs, P2.p, mx, ng, offW, offB, grad);
}

// node graph: nodes sharing an element, sorted, no self loops
static void build_graph(Ctx& c)
{
  Mpnn& M = c.mpnn;
  const HostModel& hm = c.hm;
  const int np = hm.np;
  std::vector<int> eptr(np + 1, 0), elist;
  for (int e = 0; e < hm.ne; e++)
    for (int i = 0; i < hm.nnd[e]; i++) eptr[hm.elem[8LL * e + i] + 1]++;
  for (int i = 0; i < np; i++) eptr[i + 1] += eptr[i];
  elist.resize(eptr[np]);
  {
    std::vector<int> fill(eptr.begin(), eptr.end() - 1);
    for (int e = 0; e < hm.ne; e++)
      for (int i = 0; i < hm.nnd[e]; i++) elist[fill[hm.elem[8LL * e + i]]++] = e;
  }
  std::vector<std::vector<int>> nb(np);
  parallel_for(np, [&](long long a, long long b) {
    for (long long v = a; v < b; v++) {
      std::vector<int>& L = nb[v];
      for (int k = eptr[v]; k < eptr[v + 1]; k++) {
        const int e = elist[k];
        for (int i = 0; i < hm.nnd[e]; i++) {
          const int w = hm.elem[8LL * e + i];
          if (w != v) L.push_back(w);
        }
      }
      std::sort(L.begin(), L.end());
      L.erase(std::unique(L.begin(), L.end()), L.end());
    }
  });
  std::vector<int> ptr(np + 1, 0), adj;
  for (int v = 0; v < np; v++) ptr[v + 1] = ptr[v] + (int)nb[v].size();
  adj.resize(ptr[np]);
  for (int v = 0; v < np; v++) std::copy(nb[v].begin(), nb[v].end(), adj.begin() + ptr[v]);
  M.adj_ptr.upload(ptr, c.stream);
  M.adj.upload(adj, c.stream);
  M.np = np;
  const size_t nh = (size_t)np * MC;
  M.H.alloc((ML + 1) * nh);
  M.A.alloc(ML * nh);
  M.feat.alloc((size_t)np * MF);
  M.out.alloc((size_t)np * MO);
  M.graph = true;
}

static void clear_samples(Mpnn& M)
{
  M.nsamp = 0;
  M.s_rr.clear();
  M.s_sys.clear();
  M.sys.clear();
  M.sys_version = -1;
}

// The graph follows the mesh topology; a mesh of another topology rebuilds it and drops
// the stored training samples (they belong to the old graph and its numbering).
static void need_graph(Ctx& c)
{
  if (!c.prepared || c.csr_only) fail(RMX_ERR_ARG, "learned preconditioner: needs a prepared mesh");
  Mpnn& M = c.mpnn;
  if (!M.graph || M.topo != c.topo_hash) {
    build_graph(c);
    M.topo = c.topo_hash;
    clear_samples(M);
    M.cap = 0;                                 // the next sample sizes the set again
    M.s_r.release();
    M.tn = 0;
  }
}

// features of the current state and residual r (device, length n) into M.feat; the
// squared norm of r goes to M.r2
// 1 / max_j |a_ij| per row (0 for an empty row)
__global__ void k_rowinv(int n, const int* __restrict__ rowptr, const double* __restrict__ A,
                         double* __restrict__ rinv)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m = 0.0;
  for (int p = rowptr[i]; p < rowptr[i + 1]; p++) m = fmax(m, fabs(A[p]));
  rinv[i] = m > 0.0 ? 1.0 / m : 0.0;
}

__global__ void k_mul(int n, const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ c)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = a[i] * b[i];
}

// node features of the current state into M.feat
static void features(Ctx& c)
{
  Mpnn& M = c.mpnn;
  k_feat<<<gb(M.np), 256, 0, c.stream>>>(M.np, c.U.p, c.V.p, c.S.p, c.z.p, c.cf.p, c.dflag.p,
                                         M.feat.p);
}

// q = r / row max of the current FP64 matrix, into M.q
static void scaled(Ctx& c, const double* r)
{
  Mpnn& M = c.mpnn;
  if (!c.have64) fail(RMX_ERR_ARG, "learned preconditioner: needs the FP64 matrix");
  if (M.rinv_version != c.mat_version) {
    M.rinv.alloc(c.n);
    k_rowinv<<<gb(c.n), 256, 0, c.stream>>>(c.n, c.rowptr.p, c.A64.p, M.rinv.p);
    M.rinv_version = c.mat_version;
  }
  M.q.alloc(c.n);
  k_mul<<<gb(c.n), 256, 0, c.stream>>>(c.n, r, M.rinv.p, M.q.p);
}

static void forward(Ctx& c, const float* feat, bool train);

// network outputs (D, E per node) of the current Newton system, once per system
static void trunk(Ctx& c)
{
  Mpnn& M = c.mpnn;
  if (M.trunk_version == c.mat_version) return;
  features(c);
  forward(c, M.feat.p, false);
  M.trunk_version = c.mat_version;
  M.use_net = -1;
}

static void forward(Ctx& c, const float* feat, bool train)
{
  Mpnn& M = c.mpnn;
  const int np = M.np;
  const size_t nh = (size_t)np * MC;
  if (train) { M.P.alloc(ML * nh); M.Q.alloc(nh); }
  k_in<<<gb((long long)np * MC), 256, 0, c.stream>>>(np, feat, M.theta.p, M.H.p, train ? M.Q.p : nullptr);
  for (int l = 0; l < ML; l++) {
    const uint16_t* h = M.H.p + l * nh;
    uint16_t* a = M.A.p + l * nh;
    k_agg<<<gb((long long)np * 64), 256, 0, c.stream>>>(np, M.adj_ptr.p, M.adj.p, h, a);
    k_layer_fwd<<<gb(np, 64), 256, 0, c.stream>>>(np, h, a, M.wt.p + (size_t)l * 2 * MC * MC,
                                                  M.theta.p + OFF_L0 + l * LAYER_SZ + 2 * MC * MC,
                                                  M.H.p + (l + 1) * nh, train ? M.P.p + l * nh : nullptr);
  }
  k_out<<<gb(np), 256, 0, c.stream>>>(np, M.H.p + ML * nh, M.theta.p, M.out.p);
  HIPCHECK(hipGetLastError());
}

// Training loss of one stored sample s, ||r - A n(r)||^2 / ||r||^2 with A, the row scaling
// and the network outputs (M.out, from a training forward pass of the sample's system) of its
// system: ||r - A n||^2 goes to the device double *loss, and w * d loss / d (D, E) is written
// to M.gout (add = false) or added to it (add = true).
static void sample_out_grad(Ctx& c, int s, double w, bool add, double* loss)
{
  Mpnn& M = c.mpnn;
  const int np = M.np, n = c.n;
  const MpnnSys& S = *M.sys[M.s_sys[s]];
  const double* r = M.s_r.p + (size_t)s * n;
  M.ev.alloc(3 * (size_t)n);
  double* nv = M.ev.p;
  double* res = M.ev.p + n;
  double* qs = M.ev.p + 2 * (size_t)n;
  k_mul<<<gb(n), 256, 0, c.stream>>>(n, r, S.rinv.p, qs);
  k_lin_apply<<<gb(np), 256, 0, c.stream>>>(np, c.eqno.p, M.adj_ptr.p, M.adj.p, M.out.p, qs, nv);
  k_resid_s<<<gb((long long)n * 8), 256, 0, c.stream>>>(n, c.rowptr.p, c.cols.p, S.A.p, nv, r, res);
  const double* xs[1] = { res };
  dev_dots(c, 1, xs, xs, loss);
  // d loss / d n = -2 A^T res / ||r||^2, into nv; then d loss / d (D, E)
  k_spmv_t<<<gb((long long)n * 8), 256, 0, c.stream>>>(n, M.tptr.p, M.trow.p, M.tperm.p, S.A.p, res,
                                                       -2.0 * w / M.s_rr[s], nv);
  k_lin_grad<<<gb(np), 256, 0, c.stream>>>(np, c.eqno.p, M.adj_ptr.p, M.adj.p, qs, nv, add ? 1 : 0,
                                           M.gout.p);
}

// Gradient of the weights (into M.grad) from d loss / d (D, E) in M.gout, back through the
// network of the last training forward pass (node features feat).
static void net_backward(Ctx& c, const float* feat)
{
  Mpnn& M = c.mpnn;
  const int np = M.np;
  const size_t nh = (size_t)np * MC;
  const int nchunk = gb(np, CHUNK), nw = gb(np, WCHUNK);
  M.partial.alloc(std::max<size_t>(std::max<size_t>((size_t)nchunk * (MF + 1) * MC,
                                                     (size_t)nchunk * (MC + 1) * MO),
                                    (size_t)nw * 2 * MC * MC));
  // output layer
  k_outer<uint16_t, float><<<nchunk, 256, 0, c.stream>>>(np, M.H.p + ML * nh, MC, M.gout.p, MO, M.partial.p);
  reduce_blocks(c, nchunk, (MC + 1) * MO, M.partial.p, MC, MO, OFF_WOUT, OFF_BOUT, M.grad.p);
  k_gh_out<<<gb((long long)np * MC), 256, 0, c.stream>>>(np, M.gout.p, M.theta.p, M.gh.p);
  for (int l = ML - 1; l >= 0; l--) {
    k_gp<<<2048, 256, 0, c.stream>>>((long long)nh, M.gh.p, M.P.p + l * nh, M.gp.p);
    const int off = OFF_L0 + l * LAYER_SZ;
    k_wgrad<<<nw, 256, 0, c.stream>>>(np, M.H.p + l * nh, M.A.p + l * nh, M.gp.p, M.partial.p);
    reduce_blocks(c, nw, 2 * MC * MC, M.partial.p, 2 * MC * MC, 1, off, off, M.grad.p);
    k_outer<float, uint16_t><<<nchunk, 256, 0, c.stream>>>(np, nullptr, 0, M.gp.p, MC, M.partial.p);
    reduce_blocks(c, nchunk, MC, M.partial.p, 0, MC, off, off + 2 * MC * MC, M.grad.p);
    k_layer_bwd<<<gb(np, 64), 256, 0, c.stream>>>(np, M.gp.p, M.wb.p + (size_t)l * 2 * MC * MC,
                                                  M.gh.p, M.ga.p);
    k_aggT<<<gb((long long)np * 64), 256, 0, c.stream>>>(np, M.adj_ptr.p, M.adj.p, M.ga.p, M.gh.p);
  }
  k_gq<<<2048, 256, 0, c.stream>>>((long long)nh, M.gh.p, M.Q.p);
  k_outer<float, float><<<nchunk, 256, 0, c.stream>>>(np, feat, MF, M.gh.p, MC, M.partial.p);
  reduce_blocks(c, nchunk, (MF + 1) * MC, M.partial.p, MF, MC, OFF_WIN, OFF_BIN, M.grad.p);
  HIPCHECK(hipGetLastError());
}

// training forward pass of stored system k (network outputs into M.out)
static void sys_forward(Ctx& c, int k)
{
  Mpnn& M = c.mpnn;
  const size_t nh = (size_t)M.np * MC;
  forward(c, M.sys[k]->feat.p, true);
  M.gh.alloc(nh); M.ga.alloc(nh); M.gp.alloc(nh); M.gout.alloc((size_t)M.np * MO);
}

bool mpnn_ready(Ctx& c) { return c.mpnn.weights; }

void mpnn_set_weights(Ctx& c, const float* th, int n)
{
  if (n != NPARAM) fail(RMX_ERR_ARG, "learned preconditioner: expected " + std::to_string(NPARAM) + " weights");
  Mpnn& M = c.mpnn;
  M.theta.upload(th, NPARAM, c.stream);
  M.grad.alloc(NPARAM); M.m.alloc(NPARAM); M.v.alloc(NPARAM);
  HIPCHECK(hipMemsetAsync(M.m.p, 0, sizeof(float) * NPARAM, c.stream));
  HIPCHECK(hipMemsetAsync(M.v.p, 0, sizeof(float) * NPARAM, c.stream));
  M.wt.alloc((size_t)ML * 2 * MC * MC); M.wb.alloc((size_t)ML * 2 * MC * MC);
  k_refresh<<<gb(ML * 2 * MC * MC), 256, 0, c.stream>>>(M.theta.p, M.wt.p, M.wb.p);
  HIPCHECK(hipStreamSynchronize(c.stream));
  M.step = 0;
  M.weights = true;
  M.trunk_version = -1;
}

void mpnn_get_weights(Ctx& c, float* th, int n)
{
  if (!c.mpnn.weights) fail(RMX_ERR_ARG, "learned preconditioner: no weights");
  if (n != NPARAM) fail(RMX_ERR_ARG, "learned preconditioner: expected " + std::to_string(NPARAM) + " weights");
  c.mpnn.theta.download(th, NPARAM, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
}

// z = z2 + n(r - A z2) with z2 = M2^-1 r, the two-level preconditioner's output, and n the
// network's correction, linear in its argument (device pointers)
void mpnn_apply(Ctx& c, int prec, const double* r, double* z)
{
  if (!mpnn_ready(c)) fail(RMX_ERR_ARG, "learned preconditioner not loaded");
  need_graph(c);
  Mpnn& M = c.mpnn;
  const int n = c.n;
  trunk(c);
  coarse_apply(c, prec, r, z);
  if (M.use_net == 0) return;
  M.rv.alloc(3 * (size_t)n);
  double* r2v = M.rv.p;                        // r2 = r - A z2
  double* nv = M.rv.p + n;                     // the network's correction n(r2)
  double* tv = M.rv.p + 2 * (size_t)n;         // r2 - A n
  dev_resid_p(c, prec, z, r, r2v);
  scaled(c, r2v);
  k_lin_apply<<<gb(M.np), 256, 0, c.stream>>>(M.np, c.eqno.p, M.adj_ptr.p, M.adj.p, M.out.p, M.q.p, nv);
  if (M.use_net < 0) {
    // first application of this system: keep the network only if it lowers the residual
    dev_resid_p(c, prec, nv, r2v, tv);
    M.acc.alloc(2);
    const double* xs[2] = { tv, r2v };
    dev_dots(c, 2, xs, xs, M.acc.p);
    double h[2];
    HIPCHECK(hipMemcpyAsync(h, M.acc.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    M.use_net = h[0] < h[1] ? 1 : 0;
    if (getenv("RMX_DEBUG"))
      fprintf(stderr, "[rmx] learned correction %s (residual %.3e -> %.3e)\n",
              M.use_net ? "kept" : "dropped", std::sqrt(h[1]), std::sqrt(h[0]));
    if (!M.use_net) return;
  }
  k_add_v<<<gb(n), 256, 0, c.stream>>>(n, nv, z);
  HIPCHECK(hipGetLastError());
}

// network correction alone: z = n(r) for the current system; feat (host, optional) gets
// the node features
void mpnn_forward(Ctx& c, const double* r, double* z, float* feat_out)
{
  if (!mpnn_ready(c)) fail(RMX_ERR_ARG, "learned preconditioner not loaded");
  need_graph(c);
  Mpnn& M = c.mpnn;
  M.trunk_version = -1;                        // weights may have changed
  trunk(c);
  scaled(c, r);
  k_lin_apply<<<gb(M.np), 256, 0, c.stream>>>(M.np, c.eqno.p, M.adj_ptr.p, M.adj.p, M.out.p, M.q.p, z);
  HIPCHECK(hipGetLastError());
  if (feat_out) M.feat.download(feat_out, (size_t)M.np * MF, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
}

// transpose of the system pattern: column pointers, row indices, CSR position of every
// entry (for A^T in the training loss)
static void build_transpose(Ctx& c)
{
  Mpnn& M = c.mpnn;
  const int n = c.n;
  const std::vector<int>& rp = c.hm.rowptr;
  const std::vector<int>& cl = c.hm.cols;
  std::vector<int> tptr(n + 1, 0), trow(rp[n]), tperm(rp[n]);
  for (int q = 0; q < rp[n]; q++) tptr[cl[q] + 1]++;
  for (int j = 0; j < n; j++) tptr[j + 1] += tptr[j];
  std::vector<int> fill(tptr.begin(), tptr.end() - 1);
  for (int i = 0; i < n; i++)
    for (int q = rp[i]; q < rp[i + 1]; q++) { const int k = fill[cl[q]]++; trow[k] = i; tperm[k] = q; }
  M.tptr.upload(tptr, c.stream); M.trow.upload(trow, c.stream); M.tperm.upload(tperm, c.stream);
  M.tn = n;
}

// Store a training sample: the residual r (device pointer).  The first sample of a Newton
// system (a matrix version not stored yet) also stores the system: the current FP64 matrix,
// its inverse row maxima and the node features of the current state; the later samples of
// the system share them.  capacity > 0 empties the set (samples and systems) and sizes it
// for `capacity` samples.
int mpnn_add_sample(Ctx& c, const double* r, int capacity)
{
  need_graph(c);
  if (!c.have64) fail(RMX_ERR_ARG, "learned preconditioner: a sample needs the FP64 matrix");
  Mpnn& M = c.mpnn;
  const int np = M.np, n = c.n;
  if (M.tn != n) build_transpose(c);
  if (capacity > 0) {
    clear_samples(M);
    M.cap = capacity;
    M.s_r.alloc((size_t)capacity * n);
  }
  if (M.cap == 0)
    fail(RMX_ERR_ARG, "learned preconditioner: no training set (the first sample, and the first "
                      "after a change of mesh topology or wet set, passes a capacity)");
  if (M.nsamp >= M.cap) fail(RMX_ERR_RANGE, "learned preconditioner: training set full");
  if (M.s_r.n != (size_t)M.cap * n)
    fail(RMX_ERR_ARG, "learned preconditioner: the training set holds systems of another size");
  hipStream_t st = c.stream;
  if (M.sys.empty() || M.sys_version != c.mat_version) {
    auto S = std::make_unique<MpnnSys>();
    features(c);
    S->feat.alloc((size_t)np * MF);
    S->A.alloc(c.nnz);
    S->rinv.alloc(n);
    HIPCHECK(hipMemcpyAsync(S->feat.p, M.feat.p, sizeof(float) * np * MF, hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(S->A.p, c.A64.p, sizeof(double) * c.nnz, hipMemcpyDeviceToDevice, st));
    k_rowinv<<<gb(n), 256, 0, st>>>(n, c.rowptr.p, c.A64.p, S->rinv.p);
    HIPCHECK(hipGetLastError());
    M.sys.push_back(std::move(S));
    M.sys_version = c.mat_version;
  }
  const int s = M.nsamp;
  M.lpart.alloc(1);
  HIPCHECK(hipMemcpyAsync(M.s_r.p + (size_t)s * n, r, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  double rr = 0.0;
  const double* xs[1] = { r };
  dev_dots(c, 1, xs, xs, M.lpart.p);
  HIPCHECK(hipMemcpyAsync(&rr, M.lpart.p, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  if (!(rr > 0.0)) fail(RMX_ERR_ARG, "learned preconditioner: zero residual sample");
  M.s_rr.push_back(rr);
  M.s_sys.push_back((int)M.sys.size() - 1);
  return ++M.nsamp;
}

// loss of stored sample s and its gradient (host grad, optional)
double mpnn_loss_grad(Ctx& c, int s, float* grad)
{
  Mpnn& M = c.mpnn;
  if (!M.weights) fail(RMX_ERR_ARG, "learned preconditioner: no weights");
  if (s < 0 || s >= M.nsamp) fail(RMX_ERR_ARG, "learned preconditioner: no such sample");
  need_graph(c);
  const int k = M.s_sys[s];
  sys_forward(c, k);
  M.lpart.alloc(1);
  sample_out_grad(c, s, 1.0, false, M.lpart.p);
  net_backward(c, M.sys[k]->feat.p);
  double loss = 0.0;
  HIPCHECK(hipMemcpyAsync(&loss, M.lpart.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  if (grad) M.grad.download(grad, NPARAM, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  M.trunk_version = -1;
  return loss / M.s_rr[s];
}

// Adam over the stored systems, `epochs` passes over them in order: one step per system, on
// the mean loss of its samples (the network sees only the system's state, so one forward and
// one backward pass serve all its samples).  Returns the mean sample loss of the last pass.
double mpnn_train(Ctx& c, int epochs, double lr)
{
  Mpnn& M = c.mpnn;
  if (!M.weights) fail(RMX_ERR_ARG, "learned preconditioner: no weights");
  if (M.nsamp == 0) fail(RMX_ERR_ARG, "learned preconditioner: no training samples");
  need_graph(c);
  const int nsys = (int)M.sys.size();
  std::vector<std::vector<int>> of(nsys);
  size_t most = 1;
  for (int s = 0; s < M.nsamp; s++) of[M.s_sys[s]].push_back(s);
  for (const auto& v : of) most = std::max(most, v.size());
  M.lpart.alloc(most);
  std::vector<double> lh(most);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double last = 0.0;
  for (int ep = 0; ep < epochs; ep++) {
    double sum = 0.0;
    for (int k = 0; k < nsys; k++) {
      const std::vector<int>& ss = of[k];
      if (ss.empty()) continue;
      sys_forward(c, k);
      for (size_t j = 0; j < ss.size(); j++)
        sample_out_grad(c, ss[j], 1.0 / (double)ss.size(), j > 0, M.lpart.p + j);
      net_backward(c, M.sys[k]->feat.p);
      M.step++;
      const float c1 = 1.0f - std::pow(b1, (float)M.step), c2 = 1.0f - std::pow(b2, (float)M.step);
      k_adam<<<gb(NPARAM), 256, 0, c.stream>>>(NPARAM, M.theta.p, M.grad.p, M.m.p, M.v.p, (float)lr,
                                               b1, b2, c1, c2, eps);
      k_refresh<<<gb(ML * 2 * MC * MC), 256, 0, c.stream>>>(M.theta.p, M.wt.p, M.wb.p);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(lh.data(), M.lpart.p, sizeof(double) * ss.size(), hipMemcpyDeviceToHost,
                              c.stream));
      HIPCHECK(hipStreamSynchronize(c.stream));
      for (size_t j = 0; j < ss.size(); j++) sum += lh[j] / M.s_rr[ss[j]];
    }
    last = sum / M.nsamp;
  }
  M.trunk_version = -1;
  return last;
}

}  // namespace rmx
