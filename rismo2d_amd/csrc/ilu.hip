// ILU(0) preconditioner on gfx950: factorisation and triangular sweeps
// (PRECO_ILU0::Factor / Solve, Preco_ilu0.cpp:64-186, 497-530, 958-980).
//
// An Ilu object holds one ILU(0) system: the CSR pattern, the level sets of its
// factorisation, the packed sweep layouts, and FP64 / FP32 factors.  Both instances of a
// context factor the whole matrix in the reference's equation order (its FP32 factors are
// bit-identical to the reference's); they differ in the sweeps:
//  * the reference sweeps: one workgroup walks the level sets of the whole lower / upper
//    factor (the plain FP64 path);
//  * strip sweeps: rows are cut into contiguous ranges (the frontal order runs along the
//    channel, so a range is a strip of cross-sections) and the factor entries that couple
//    two strips are left out of the sweeps, so each strip is swept by its own workgroup.
//    The pivots stay those of the full factorisation (factoring the strips separately
//    loses the upstream couplings at each strip start and leaves near-zero pivots).
//    It preconditions the FP32 solve inside the FP64 refinement, where the FP64 residual,
//    not the preconditioner, decides the answer.
//
// Factorisation: row-wise, one wave per row, the rows of one level in one launch; row i
// eliminates its lower entries in increasing column order, so every factor entry gets its
// updates in the reference's order and, with FMA contraction off, the reference's
// rounding.  Multipliers are stored negated as in the reference.
//
// Sweeps: one workgroup per strip walks its level sets.  Rows are stored in processing
// order with each row's off-diagonal entries contiguous and columns replaced by
// processing positions.  Solved values go to an LDS ring (slot = position & (RING-1)) and
// to a position-ordered global copy; a level reads a dependency from the ring while its
// slot has not been reused (position >= P + nr - RING for a level of nr rows starting at
// P) and from the global copy otherwise, which is safe because a full barrier (global
// release / acquire) is taken whenever the position crosses a multiple of FENCE, and any
// value that has left the ring is older than the last such barrier (nr <= RING/4).
// Wider levels read everything from the global copy between full barriers.  Other levels
// are separated by LDS-only barriers, so the next level's rows, loaded into registers
// while the current level finishes, stay in flight across the barrier.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "rmx_internal.h"

namespace rmx {

constexpr int FW = 128;         // maximum row width handled by the factorisation
constexpr int PW = 6;           // entries per lane held in registers (more are read on)
constexpr int SG = 8;           // lanes per row in the sweeps
constexpr int SKR = 2;          // row-rounds per level held in registers
constexpr int LPC = 1024;       // level offsets staged in LDS at a time
constexpr int FENCE = 2048;     // positions between full barriers
constexpr int WIDE = RING / 4;  // levels wider than this use the global copy only
constexpr int STAGE_THREADS = 512;                 // staged sweep: threads per workgroup
constexpr int STAGE_ROWS = STAGE_THREADS / 4;      // rows per round (4 lanes per row)

// ---------------------------------------------------------------------------------------
// factorisation

constexpr int UCAP = 2048;      // entries of the pivot rows staged in LDS per row
constexpr int UPL = UCAP / 64;  // staged entries per lane

// One row of one level (one wave): row i eliminates its lower entries k in increasing
// column order (IKJ).  Per-pattern metadata keeps the dependent memory rounds to three: the
// row's descriptor {i, start, width, lower count}; the row (values, sorted columns) with
// the start and length of every pivot row; then all pivot rows at once, staged in LDS with
// independent loads, every staged entry getting its target position in row i (a binary
// search over row i's sorted columns, done in parallel).  The sequential loop over k is
// then one division and one LDS update per entry.  The updates are the reference's, in its
// order, with FMA contraction off.  Rows whose pivot rows hold more than UCAP entries
// search and read from global memory inside the loop instead.
// Symbolic plan of the factorisation (once per pattern, all rows in one launch, one wave
// per row): every entry u of a row's concatenated pivot rows (the entries after each pivot
// row's diagonal, in sorted lower-column order of the row) gets its pivot index t (low
// byte) and its target position in the row (high byte; 255: not in the row, or a column at
// or left of the pivot).
__global__ void __launch_bounds__(64) k_ilu_plan(const int4* __restrict__ desc, int cnt,
                                                 const int* __restrict__ cols,
                                                 const int* __restrict__ srt_cols,
                                                 const unsigned char* __restrict__ srt_pos,
                                                 const int2* __restrict__ piv,
                                                 const long long* __restrict__ pbase,
                                                 unsigned short* __restrict__ plan)
{
  __shared__ int s_col[FW], s_kb[FW], s_off[FW + 1];
  __shared__ unsigned char s_pos[FW];
  const int lane = threadIdx.x;
  const int w = blockIdx.x;
  if (w >= cnt) return;
  const int4 d = desc[w];
  const int p0 = d.y, wi = d.z, nk = d.w;
  for (int t = lane; t < wi; t += 64) {
    s_col[t] = srt_cols[p0 + t];
    s_pos[t] = srt_pos[p0 + t];
  }
  int carry = 0;
  for (int t0 = 0; t0 < nk; t0 += 64) {
    const int t = t0 + lane;
    int len = 0;
    if (t < nk) {
      const int2 pv = piv[p0 + t];
      s_kb[t] = pv.x;
      len = pv.y;
    }
    int incl = len;
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (t < nk) s_off[t + 1] = carry + incl;
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) s_off[0] = 0;
  __syncthreads();
  const int total = s_off[nk];
  const long long pb = pbase[w];
  for (int u = lane; u < total; u += 64) {
    int lo = 0, hi = nk - 1;                // pivot t with s_off[t] <= u < s_off[t+1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_off[mid] <= u) lo = mid; else hi = mid - 1;
    }
    const int k = s_col[lo];
    const int cc = cols[s_kb[lo] + 1 + (u - s_off[lo])];
    int pos = 255;
    if (cc > k) {
      int a = 0, b = wi - 1;
      while (a <= b) {
        const int mid = (a + b) >> 1;
        const int v = s_col[mid];
        if (v == cc) { pos = s_pos[mid]; break; }
        if (v < cc) a = mid + 1; else b = mid - 1;
      }
    }
    plan[pb + u] = (unsigned short)(lo | (pos << 8));
  }
}

// Staged pivot-row entries of one row, in rounds of 64 (entry u = lane + 64 m) grouped in
// chunks of PCH rounds: a chunk is taken when the row has entries in it (a wave-uniform
// test), and inside a chunk no lane branches around a load (lanes past the last entry read
// it again), so a chunk's loads are all issued before the first wait.  plan_offsets: the
// target position of every entry and the offset (from element base) of its value in its
// pivot row; stage_entries: the loaded values and targets into LDS.
constexpr int PCH = 8;
__device__ __forceinline__ void plan_offsets(const unsigned short* __restrict__ plan, int total,
                                             const int* s_kb, const int* s_off, int base,
                                             unsigned* offs, unsigned char* pp)
{
  const int lane = threadIdx.x;
  unsigned short tp[UPL];
#pragma unroll
  for (int c = 0; c < UPL / PCH; c++)
    if (c * PCH * 64 < total) {
#pragma unroll
      for (int m = c * PCH; m < (c + 1) * PCH; m++) tp[m] = plan[min(lane + 64 * m, total - 1)];
#pragma unroll
      for (int m = c * PCH; m < (c + 1) * PCH; m++) {
        const int u = min(lane + 64 * m, total - 1);
        const int t = tp[m] & 0xFF;
        pp[m] = (unsigned char)(tp[m] >> 8);
        offs[m] = (unsigned)(s_kb[t] - base + 1 + (u - s_off[t]));
      }
    }
}

template <class VT>
__device__ __forceinline__ void stage_entries(int total, const unsigned* offs,
                                              const unsigned char* pp, const VT* uv,
                                              unsigned char* s_upos, VT* s_uval)
{
  const int lane = threadIdx.x;
#pragma unroll
  for (int c = 0; c < UPL / PCH; c++)
    if (c * PCH * 64 < total) {
#pragma unroll
      for (int m = c * PCH; m < (c + 1) * PCH; m++) {
        const int u = lane + 64 * m;
        if (u < total) {
          s_upos[u] = pp[m];
          s_uval[u] = uv[m];
        }
      }
    }
}

template <class VT>
__global__ void __launch_bounds__(64) k_ilu_level(const int4* __restrict__ desc, int cnt,
                                                  const int* __restrict__ cols,
                                                  const int* __restrict__ srt_cols,
                                                  const unsigned char* __restrict__ srt_pos,
                                                  const int2* __restrict__ piv,
                                                  const long long* __restrict__ pbase,
                                                  const unsigned short* __restrict__ plan,
                                                  VT* __restrict__ LU, int* __restrict__ err,
                                                  const int4* __restrict__ desc_next,
                                                  int cnt_next,
                                                  const long long* __restrict__ pbase_next,
                                                  long long plan_n)
{
#pragma clang fp contract(off)
  __shared__ VT s_val[FW];
  __shared__ int s_col[FW];
  __shared__ unsigned char s_pos[FW];
  __shared__ int s_kb[FW], s_off[FW + 1];
  __shared__ VT s_dk[FW];
  __shared__ unsigned char s_upos[UCAP];
  __shared__ VT s_uval[UCAP];
  const int lane = threadIdx.x;
  const int w = blockIdx.x;
  if (w >= cnt) return;                 // uniform per block
  const int4 d = desc[w];
  const int i = d.x, p0 = d.y, wi = d.z, nk = d.w;
  // the next level's row at this index: its descriptor now, its static data (and initial
  // values, which no other row writes) touched after this row's own loads are issued, so
  // the next launch finds them in the memory-side cache
  const bool pre = w < cnt_next;
  int4 dn = make_int4(0, 0, 0, 0);
  long long pbn = 0;
  if (pre) { dn = desc_next[w]; pbn = pbase_next[w]; }
  for (int t = lane; t < wi; t += 64) {
    s_val[t] = LU[p0 + t];
    s_col[t] = srt_cols[p0 + t];
    s_pos[t] = srt_pos[p0 + t];
  }
  // pivot rows (the first nk sorted columns): start and entries after the diagonal;
  // offsets of their staged entries by an inclusive wave scan
  int carry = 0;
  for (int t0 = 0; t0 < nk; t0 += 64) {
    const int t = t0 + lane;
    int len = 0;
    if (t < nk) {
      const int2 pv = piv[p0 + t];
      s_kb[t] = pv.x;
      len = pv.y;
    }
    int incl = len;
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (t < nk) s_off[t + 1] = carry + incl;
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) s_off[0] = 0;
  __syncthreads();
  // target position in row i of a column cc of pivot row k (255: not in row i, or cc <= k)
  auto target = [&](int cc, int k) -> int {
    if (cc <= k) return 255;
    int lo = 0, hi = wi - 1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int v = s_col[mid];
      if (v == cc) return s_pos[mid];
      if (v < cc) lo = mid + 1; else hi = mid - 1;
    }
    return 255;
  };
  const int total = s_off[nk];
  const bool staged = total <= UCAP;
  for (int t = lane; t < nk; t += 64) s_dk[t] = LU[s_kb[t]];
  if (staged) {                         // pivot-row values, targets from the plan
    const long long pb = pbase[w];
    unsigned offs[UPL];
    unsigned char pp[UPL];
    plan_offsets(plan + pb, total, s_kb, s_off, 0, offs, pp);
    VT uv[UPL];
#pragma unroll
    for (int c = 0; c < UPL / PCH; c++)
      if (c * PCH * 64 < total) {
#pragma unroll
        for (int m = c * PCH; m < (c + 1) * PCH; m++) uv[m] = LU[offs[m]];
      }
    stage_entries(total, offs, pp, uv, s_upos, s_uval);
  }
  unsigned int sink = 0;
  if (pre) {
    for (int t = lane; t < dn.z; t += 64)
      sink += (unsigned int)srt_cols[dn.y + t] + (unsigned int)srt_pos[dn.y + t] +
              (unsigned int)(LU[dn.y + t] != (VT)0);
    for (int t = lane; t < dn.w; t += 64) sink += (unsigned int)piv[dn.y + t].x;
    for (long long u = pbn + 128LL * lane; u < pbn + UCAP && u < plan_n; u += 128LL * 64)
      sink += (unsigned int)plan[u];
  }
  __syncthreads();
  if (staged) {
    // pivot t's diagonal position and value and this lane's (at most two) entries of its
    // pivot row (target, value) are read one pivot ahead, so that a pivot's chain is the
    // read of the two row values, the division and the update
    auto load_piv = [&](int t, int& pik, VT& dk, int& pa, VT& ua, int& pb, VT& ub_) {
      pik = s_pos[t];
      dk = s_dk[t];
      const int lo = s_off[t], hi = s_off[t + 1];
      const int e0 = lo + lane, e1 = e0 + 64;
      pa = 255; pb = 255; ua = (VT)0; ub_ = (VT)0;
      if (e0 < hi) { pa = s_upos[e0]; ua = s_uval[e0]; }
      if (e1 < hi) { pb = s_upos[e1]; ub_ = s_uval[e1]; }
    };
    int pik = 0, pa = 255, pb = 255;
    VT dk = (VT)1, ua = (VT)0, ub_ = (VT)0;
    if (nk > 0) load_piv(0, pik, dk, pa, ua, pb, ub_);
    for (int t = 0; t < nk; t++) {
      int pikn = 0, pan = 255, pbn = 255;
      VT dkn = (VT)1, uan = (VT)0, ubn = (VT)0;
      if (t + 1 < nk) load_piv(t + 1, pikn, dkn, pan, uan, pbn, ubn);
      const VT a = s_val[pik];
      const VT ca = pa != 255 ? s_val[pa] : (VT)0;
      const VT cb = pb != 255 ? s_val[pb] : (VT)0;
      const VT f = -a / dk;                // every lane, from the same LDS values
      if (pa != 255) { const VT prod = f * ua; s_val[pa] = ca + prod; }
      if (pb != 255) { const VT prod = f * ub_; s_val[pb] = cb + prod; }
      __syncthreads();                    // the updates before the next pivot
      if (lane == 0) s_val[pik] = f;      // the stored multiplier (no later pivot reads it)
      pik = pikn; dk = dkn; pa = pan; ua = uan; pb = pbn; ub_ = ubn;
    }
  } else {
    for (int t = 0; t < nk; t++) {
      const int k = s_col[t];
      const int pik = s_pos[t];
      const VT f = -s_val[pik] / s_dk[t];
      const int ua = s_off[t], ub = s_off[t + 1];
      for (int e = lane; e < ub - ua; e += 64) {
        const int q = s_kb[t] + 1 + e;
        const int pos = target(cols[q], k);
        if (pos != 255) {
          const VT prod = f * LU[q];
          s_val[pos] = s_val[pos] + prod;
        }
      }
      __syncthreads();                    // the stored multiplier and updates before the next k
      if (lane == 0) s_val[pik] = f;
    }
  }
  __syncthreads();
  for (int t = lane; t < wi; t += 64) LU[p0 + t] = s_val[t];
  if (lane == 0 && fabs((double)s_val[0]) < 1.0e-20) atomicOr(err, 1);
  if (sink == 0x9E3779B9u && cnt < 0) err[1] = (int)sink;   // keeps the touches (never taken)
}

// ---------------------------------------------------------------------------------------
// dataflow factorisation: one persistent launch instead of one launch per level

// Loads and stores of factor values that other waves of the same launch write or read are
// agent-scope atomics (sc1: coherent across the XCDs' L2 caches), and so are the row
// completion flags.  A row publishes its values, waits until they are performed
// (vmcnt(0): the stores are acknowledged at agent scope) and then stores its flag; a
// consumer issues its loads of a pivot row only after its flag load has returned the
// current epoch.  No cache-wide write-back or invalidate is needed.
__device__ __forceinline__ int ld_flag(const int* p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_flag(int* p, int v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Factor values through a buffer resource based at a wave-uniform element (32-bit byte
// offsets from it), with the sc1 cache policy: plain loads and stores, so the compiler
// batches them under one wait, and coherent at agent scope.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t sc1_rsrc(const void* base)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, -1, 0x00020000);
}
__device__ __forceinline__ float ld_sc1(__amdgpu_buffer_rsrc_t r, unsigned e, float)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, e * 4u, 0, 16));
}
__device__ __forceinline__ double ld_sc1(__amdgpu_buffer_rsrc_t r, unsigned e, double)
{
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, e * 8u, 0, 16));
}
__device__ __forceinline__ void st_sc1(__amdgpu_buffer_rsrc_t r, unsigned e, float v)
{
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, e * 4u, 0, 16);
}
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_sc1(__amdgpu_buffer_rsrc_t r, unsigned e, double v)
{
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, e * 8u, 0, 16);
}

enum { FLOW_COUNTER = 0, FLOW_ERR = 1, FLOW_ABORT = 2, FLOW_BADROW = 3 };

// Ordering of LDS accesses inside one wave: a one-wave workgroup needs no barrier (a wave's
// LDS operations complete in order); the fences keep the compiler from moving accesses
// across it.
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Waves take rows in level order from a counter (one device-scope fetch-add per row, about
// 12 ns on this part).  A pivot row is earlier in that order than the rows that need it, so
// it has been taken by a running wave and every wait ends; no wave has to be resident
// before it takes work.  Per row: the row's own values, its sorted columns and plan are
// read with plain loads (no other wave writes them).  The pivots are then eliminated in
// increasing column order, as the reference does (update order kept, FMA contraction off),
// progressively: the run of pivots from the next one whose rows are done (one flag load
// per pivot) is staged and eliminated; when the next pivot's row is not done, one lane
// polls its flag (bounded by spin_limit polls, after which the row sets the abort flag
// that ends every wave's waits).  The rows finished last are usually the last pivots, so
// most of a row's work is done before its last pivot row arrives.  The finished row is
// stored and its flag set to the factorisation's epoch.  Rows whose pivot rows hold more
// entries than the LDS staging first wait for all their pivots and read the entries from
// memory.
template <class VT>
__global__ void __launch_bounds__(64) k_ilu_flow(const int4* __restrict__ desc, int n,
                                                 const int* __restrict__ cols,
                                                 const int* __restrict__ srt_cols,
                                                 const unsigned char* __restrict__ srt_pos,
                                                 const int2* __restrict__ piv,
                                                 const long long* __restrict__ pbase,
                                                 const unsigned short* __restrict__ plan,
                                                 VT* __restrict__ LU, int* __restrict__ done,
                                                 int epoch, int* __restrict__ state,
                                                 int spin_limit, int skip_upto)
{
#pragma clang fp contract(off)
  __shared__ VT s_val[FW];
  __shared__ int s_col[FW];
  __shared__ unsigned char s_pos[FW];
  __shared__ int s_kb[FW], s_off[FW + 1];
  __shared__ VT s_dk[FW];
  __shared__ unsigned char s_upos[UCAP];
  __shared__ VT s_uval[UCAP];
  const int lane = threadIdx.x;
  int polls = 0;
  // one lane polls the flag of row k until it holds the epoch; false: aborted
  auto wait_row = [&](int k) -> bool {
    for (;;) {
      int v = 0;
      if (lane == 0) v = ld_flag(done + k);
      if (__builtin_amdgcn_readfirstlane(v) == epoch) return true;
      if ((++polls & 31) == 0) {
        int ab = 0;
        if (lane == 0) ab = ld_flag(state + FLOW_ABORT);
        if (__builtin_amdgcn_readfirstlane(ab) != 0 || polls > spin_limit) return false;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  };
  for (;;) {
    int w = 0;
    if (lane == 0) w = __hip_atomic_fetch_add(state + FLOW_COUNTER, 1, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
    w = __builtin_amdgcn_readfirstlane(w);
    if (w >= n) break;
    const int4 d = desc[w];
    const int i = d.x, p0 = d.y, wi = d.z, nk = d.w;
    if (i <= skip_upto) continue;                  // final from an earlier pass
    for (int t = lane; t < wi; t += 64) {
      s_val[t] = LU[p0 + t];
      s_col[t] = srt_cols[p0 + t];
      s_pos[t] = srt_pos[p0 + t];
    }
    int carry = 0;
    for (int t0 = 0; t0 < nk; t0 += 64) {
      const int t = t0 + lane;
      int len = 0;
      if (t < nk) {
        const int2 pv = piv[p0 + t];
        s_kb[t] = pv.x;
        len = pv.y;
      }
      int incl = len;
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
      }
      if (t < nk) s_off[t + 1] = carry + incl;
      carry += __shfl(incl, 63, 64);
    }
    if (lane == 0) s_off[0] = 0;
    wave_sync();
    const int total = s_off[nk];
    const bool staged = total <= UCAP;
    // pivot rows lie within one band of rows: offsets from the first fit 32 bits
    const int kb0 = nk > 0 ? __builtin_amdgcn_readfirstlane(s_kb[0]) : p0;
    const __amdgpu_buffer_rsrc_t rp = sc1_rsrc(LU + kb0);
    bool aborted = false;
    polls = 0;
    if (staged) {
      unsigned offs[UPL];
      unsigned char pp[UPL];
      plan_offsets(plan + pbase[w], total, s_kb, s_off, kb0, offs, pp);
      int ta = 0;
      bool waited = false;                        // pivot ta's row was just seen done
      while (ta < nk) {
        // tb: the first pivot from ta on whose row is not done (after a wait: ta alone,
        // whose flag was read last, so its entries load without another flag round trip)
        int tb = ta + 1;
        if (!waited) {
          tb = nk;
          for (int t0 = ta; t0 < nk; t0 += 64) {
            const int t = t0 + lane;
            const bool ok = t >= nk || ld_flag(done + s_col[t]) == epoch;
            const unsigned long long pend = __ballot(!ok);
            if (pend) { tb = t0 + (int)__builtin_ctzll(pend); break; }
          }
          if (tb == ta) {
            if (!wait_row(s_col[ta])) { aborted = true; break; }
            waited = true;
            continue;
          }
        }
        waited = false;
        // pivots [ta, tb) are done: loads of their rows only after the flags (an acquire
        // at the IR level: nothing below is hoisted above the flag loads; no instruction)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (int t = ta + lane; t < tb; t += 64) s_dk[t] = ld_sc1(rp, (unsigned)(s_kb[t] - kb0), VT());
        const int ua = s_off[ta], ub = s_off[tb];
        VT uv[UPL];
#pragma unroll
        for (int c = 0; c < UPL / PCH; c++)
          if (c * PCH * 64 < ub && (c + 1) * PCH * 64 > ua) {
#pragma unroll
            for (int m = c * PCH; m < (c + 1) * PCH; m++) uv[m] = ld_sc1(rp, offs[m], VT());
          }
#pragma unroll
        for (int c = 0; c < UPL / PCH; c++)
          if (c * PCH * 64 < ub && (c + 1) * PCH * 64 > ua) {
#pragma unroll
            for (int m = c * PCH; m < (c + 1) * PCH; m++) {
              const int u = lane + 64 * m;
              if (u >= ua && u < ub) {
                s_upos[u] = pp[m];
                s_uval[u] = uv[m];
              }
            }
          }
        wave_sync();
        // eliminate pivots [ta, tb); pivot t's diagonal position and value and this lane's
        // (at most two) entries of its row are read one pivot ahead
        auto load_piv = [&](int t, int& pik, VT& dk, int& pa, VT& ua_, int& pb, VT& ub_) {
          pik = s_pos[t];
          dk = s_dk[t];
          const int lo = s_off[t], hi = s_off[t + 1];
          const int e0 = lo + lane, e1 = e0 + 64;
          pa = 255; pb = 255; ua_ = (VT)0; ub_ = (VT)0;
          if (e0 < hi) { pa = s_upos[e0]; ua_ = s_uval[e0]; }
          if (e1 < hi) { pb = s_upos[e1]; ub_ = s_uval[e1]; }
        };
        int pik = 0, pa = 255, pb = 255;
        VT dk = (VT)1, va = (VT)0, vb = (VT)0;
        load_piv(ta, pik, dk, pa, va, pb, vb);
        for (int t = ta; t < tb; t++) {
          int pikn = 0, pan = 255, pbn = 255;
          VT dkn = (VT)1, van = (VT)0, vbn = (VT)0;
          if (t + 1 < tb) load_piv(t + 1, pikn, dkn, pan, van, pbn, vbn);
          const VT a = s_val[pik];
          const VT ca = pa != 255 ? s_val[pa] : (VT)0;
          const VT cb = pb != 255 ? s_val[pb] : (VT)0;
          const VT f = -a / dk;
          if (pa != 255) { const VT prod = f * va; s_val[pa] = ca + prod; }
          if (pb != 255) { const VT prod = f * vb; s_val[pb] = cb + prod; }
          wave_sync();
          if (lane == 0) s_val[pik] = f;           // the stored multiplier
          wave_sync();
          pik = pikn; dk = dkn; pa = pan; va = van; pb = pbn; vb = vbn;
        }
        ta = tb;
      }
    } else {
      for (int t0 = 0; t0 < nk && !aborted; t0 += 64) {
        const int t = t0 + lane;
        const bool ok = t >= nk || ld_flag(done + s_col[t]) == epoch;
        unsigned long long pending = __ballot(!ok);
        while (pending && !aborted) {
          if (!wait_row(s_col[t0 + (int)__builtin_ctzll(pending)])) aborted = true;
          pending &= pending - 1;
        }
      }
      if (!aborted) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (int t = lane; t < nk; t += 64) s_dk[t] = ld_sc1(rp, (unsigned)(s_kb[t] - kb0), VT());
        wave_sync();
        for (int t = 0; t < nk; t++) {
          const int k = s_col[t];
          const int pik = s_pos[t];
          const VT f = -s_val[pik] / s_dk[t];
          const int ua = s_off[t], ub = s_off[t + 1];
          for (int e = lane; e < ub - ua; e += 64) {
            const int q = s_kb[t] + 1 + e;
            const int cc = cols[q];
            int pos = 255;
            if (cc > k) {
              int lo = 0, hi = wi - 1;
              while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const int v = s_col[mid];
                if (v == cc) { pos = s_pos[mid]; break; }
                if (v < cc) lo = mid + 1; else hi = mid - 1;
              }
            }
            if (pos != 255) {
              const VT prod = f * ld_sc1(rp, (unsigned)(q - kb0), VT());
              s_val[pos] = s_val[pos] + prod;
            }
          }
          wave_sync();
          if (lane == 0) s_val[pik] = f;
          wave_sync();
        }
      }
    }
    if (aborted) {
      if (lane == 0) st_flag(state + FLOW_ABORT, 1);
      break;
    }
    wave_sync();
    const __amdgpu_buffer_rsrc_t ro = sc1_rsrc(LU + p0);
    for (int t = lane; t < wi; t += 64) st_sc1(ro, (unsigned)t, s_val[t]);
    const bool bad = fabs((double)s_val[0]) < 1.0e-20;
    __builtin_amdgcn_s_waitcnt(0x0F70);            // vmcnt(0): the row is performed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == 0) {
      st_flag(done + i, epoch);
      if (bad) {
        atomicOr(state + FLOW_ERR, 1);
        atomicMin(state + FLOW_BADROW, i);
      }
    }
    wave_sync();                                   // LDS reused by the next row
  }
}

// Pivot repair of row i (PRECO_ILU0::Factor, Preco_ilu0.cpp:95-135), every row before i
// final: among the rows eq > i of row i's pattern that hold column i, the one whose entry
// in column i is largest in magnitude (the first in row i's order on a tie) after its
// elimination by its pivot rows before i, is added to row i on row i's pattern.  The
// candidates' partial rows start from the matrix A (their factor rows may hold values of
// the failed pass).  One wave; state[FLOW_ERR] = 2 when no row qualifies.
__device__ __forceinline__ int row_pos(const int* __restrict__ srt_cols,
                                       const unsigned char* __restrict__ srt_pos, int q0, int w,
                                       int c)
{
  int lo = 0, hi = w - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int v = srt_cols[q0 + mid];
    if (v == c) return srt_pos[q0 + mid];
    if (v < c) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

template <class VT>
__global__ void __launch_bounds__(64) k_ilu_repair(int i, const int* __restrict__ rp,
                                                   const int* __restrict__ cols,
                                                   const int* __restrict__ srt_cols,
                                                   const unsigned char* __restrict__ srt_pos,
                                                   const VT* __restrict__ A, VT* __restrict__ LU,
                                                   int* __restrict__ state)
{
#pragma clang fp contract(off)
  __shared__ VT s_v[FW], s_best[FW];
  const int lane = threadIdx.x;
  const int p0 = rp[i], wi = rp[i + 1] - p0;
  double bmax = 0.0;
  int best = -1, bq0 = 0, bw = 0;
  for (int j = 1; j < wi; j++) {
    const int eq = cols[p0 + j];
    if (eq <= i) continue;
    const int q0 = rp[eq], wq = rp[eq + 1] - q0;
    const int pos_i = row_pos(srt_cols, srt_pos, q0, wq, i);
    if (pos_i < 0) continue;
    for (int t = lane; t < wq; t += 64) s_v[t] = A[q0 + t];
    __syncthreads();
    // elimination by the pivot rows p < i of row eq, ascending
    for (int sidx = 0; sidx < wq; sidx++) {
      const int p = srt_cols[q0 + sidx];
      if (p >= i) break;
      const int pp = srt_pos[q0 + sidx];
      const int k0 = rp[p], wk = rp[p + 1] - k0;
      const VT f = -s_v[pp] / LU[k0];
      __syncthreads();
      for (int e = lane; e < wk - 1; e += 64) {
        const int q = k0 + 1 + e;
        const int c = cols[q];
        if (c > p) {
          const int pos = row_pos(srt_cols, srt_pos, q0, wq, c);
          if (pos >= 0) {
            const VT prod = f * LU[q];
            s_v[pos] = s_v[pos] + prod;
          }
        }
      }
      __syncthreads();
      if (lane == 0) s_v[pp] = f;
      __syncthreads();
    }
    const double m = fabs((double)s_v[pos_i]);
    if (m > bmax) {
      bmax = m;
      best = eq; bq0 = q0; bw = wq;
      for (int t = lane; t < wq; t += 64) s_best[t] = s_v[t];
    }
    __syncthreads();
  }
  if (best < 0) {
    if (lane == 0) state[FLOW_ERR] = 2;
    return;
  }
  for (int t = lane; t < wi; t += 64) {
    const int pos = row_pos(srt_cols, srt_pos, bq0, bw, cols[p0 + t]);
    if (pos >= 0) LU[p0 + t] = LU[p0 + t] + s_best[pos];
  }
}

// Completion flags of the rows after `from` cleared (their factorisation runs again).
__global__ void k_flags_clear(int from, int n, int* __restrict__ done)
{
  const int r = from + blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) done[r] = 0;
}

template <class VT>
__global__ void k_pack(long long ne, const int* __restrict__ psrc, const VT* __restrict__ LU,
                       VT* __restrict__ pval)
{
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ne;
       t += (long long)gridDim.x * blockDim.x)
    pval[t] = psrc[t] >= 0 ? LU[psrc[t]] : (VT)0;       // negative: a zero (padding) entry
}

// Reciprocal diagonal in FP64 (1 for a negative source: an absent row).
template <class VT>
__global__ void k_pack_rcp(long long ne, const int* __restrict__ psrc, const VT* __restrict__ LU,
                           double* __restrict__ out)
{
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ne;
       t += (long long)gridDim.x * blockDim.x)
    out[t] = psrc[t] >= 0 ? 1.0 / (double)LU[psrc[t]] : 1.0;
}

template <class VT>
__global__ void k_pack_diag(int n, const int* __restrict__ prow, const int* __restrict__ rowptr,
                            const VT* __restrict__ LU, VT* __restrict__ pdiag)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) pdiag[r] = LU[rowptr[prow[r]]];
}

// ---------------------------------------------------------------------------------------
// sweeps

__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ double ring_or_global(const double* ring, const double* xp, int c, int lim)
{
  double v;
  if (c >= lim) v = ring[c & (RING - 1)];
  else v = __builtin_nontemporal_load(xp + c);
  return v;
}

// One lane's share of a row's sum over entries q0, q0 + G, ... below q1, each solved value
// from the LDS ring or, when it has left the ring, from the global copy (out of line: the
// sweeps take it only for long rows, wide levels and old dependencies).
template <class VT, int G>
__device__ __noinline__ double lane_sum(const VT* __restrict__ pval, const int* __restrict__ pcol,
                                        const double* ring, const double* __restrict__ xp,
                                        int q0, int q1, int lim)
{
  double s = 0.0;
  for (int q = q0; q < q1; q += G) {
    const int c = pcol[q];
    s += (double)pval[q] * (c >= lim ? ring[c & (RING - 1)] : xp[c]);
  }
  return s;
}

// Block b of the sweep: levels [blev[b], blev[b+1]) of the concatenated level table, rows
// at positions [bpos[b], bpos[b+1]).  Level offsets, columns and ring slots are local to
// the block; each block table has one extra sentinel level offset.
// The right-hand side of position r is gsrc[gidx[r]] (the input vector gathered into
// processing order on the fly); with oidx the solution of position r is also written to
// odst[oidx[r]] (scattered back to equation order).
template <class VT, bool LOWER, int G, int KR>
__global__ void __launch_bounds__(1024) k_sweep(const int* __restrict__ blev,
                                                const int* __restrict__ bpos,
                                                const int* __restrict__ lev_ptr_all,
                                                const int* __restrict__ prs_all,
                                                const int* __restrict__ pcol,
                                                const VT* __restrict__ pval,
                                                const VT* __restrict__ pdiag_all,
                                                const int* __restrict__ gidx_all,
                                                const double* __restrict__ gsrc,
                                                const int* __restrict__ oidx_all,
                                                double* __restrict__ odst,
                                                double* __restrict__ xp_all)
{
  __shared__ double ring[RING];
  __shared__ int s_lp[LPC + 2];
  const int b = blockIdx.x;
  const int lb = blev[b];
  const int nlev = blev[b + 1] - lb;
  const int* lev_ptr = lev_ptr_all + lb + 2 * b;   // nlev + 2 local offsets
  const int base = bpos[b];
  const int* prs = prs_all + base;
  const VT* pdiag = LOWER ? nullptr : pdiag_all + base;
  const int* gidx = gidx_all + base;
  const int* oidx = oidx_all ? oidx_all + base : nullptr;
  double* xp = xp_all + base;
  const int t = threadIdx.x;
  const int grp = t / G, lane = t % G, ngr = blockDim.x / G;
  auto load_lp = [&](int l0) {
    for (int k = t; k < LPC + 2; k += blockDim.x) s_lp[k] = lev_ptr[min(l0 + k, nlev + 1)];
  };
  load_lp(0);
  __syncthreads();
  if (nlev <= 0) return;
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
        inr[k] = gsrc[gidx[r]];
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
    // positions >= lim are still in the ring; a wide level reads the global copy only
    const int lim = nr > WIDE ? INT_MAX : P + nr - RING;
    const int rounds = (nr + ngr - 1) / ngr;
    for (int k = 0; k < rounds; k++) {
      const int g = grp + k * ngr;
      const bool act = g < nr;
      double s = 0.0;
      double in = 0.0;
      VT dg = (VT)1;
      if (act) {
        if (k < KR) {
          const int kk = k < KR ? k : 0;
          in = inr[kk];
          dg = dgr[kk];
          bool old = false;
#pragma unroll
          for (int j = 0; j < PW; j++) {
            const int c = cr[kk][j];
            if (c >= 0) {
              old |= c < lim;
              s += (double)vr[kk][j] * ring[c & (RING - 1)];
            }
          }
          if (old) {                         // a dependency has left the ring: redo the row
            s = 0.0;
            for (int j = 0; j < PW; j++) {
              const int c = cr[kk][j];
              if (c >= 0) s += (double)vr[kk][j] * ring_or_global(ring, xp, c, lim);
            }
          }
          for (int q = qar[kk] + lane + PW * G; q < qbr[kk]; q += G)
            s += (double)pval[q] * ring_or_global(ring, xp, pcol[q], lim);
        } else {
          const int r = P + g;
          in = gsrc[gidx[r]];
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
        if (oidx) odst[oidx[P + g]] = xi;
      }
    }
    const int Pn = P + nr;
    bool full = Pn / FENCE != P / FENCE || nr > WIDE;
    if (l + 1 < nlev) {
      if (l + 2 - lbase > LPC) {            // next chunk of level offsets (rare)
        __syncthreads();
        lbase = l + 1;
        load_lp(lbase);
        __syncthreads();
      }
      P = s_lp[l + 1 - lbase];
      nr = s_lp[l + 2 - lbase] - P;
      full |= nr > WIDE;                    // a wide level starts after a full barrier
      fetch(P, nr);                         // next level's registers, in flight across the barrier
    }
    if (full) __syncthreads();
    else lds_barrier();
  }
}

// Sweep with a two-stage load pipeline and no global stores inside a level.  Level l
// computes from registers loaded during level l - 1 (stage 2: columns, values, right-hand
// side, diagonal); during level l the stage-2 loads of level l + 1 and the stage-1 loads
// of level l + 2 (row bounds, right-hand-side index) are issued, and every level starts by
// waiting for what the previous level issued, so no load is waited for in the level that
// issues it.  Results stay in the LDS ring and are flushed to the global copy (and the
// scatter target) at the full barriers: every FENCE positions, around wide levels and at
// the end.  A wide level (more than WIDE rows) reads and writes the global copy directly.
// With aidx, the scatter adds aval[aidx[e]] to the value written to odst[e] (the coarse
// prolongation P zc of the two-level preconditioner, folded into the write).
template <class VT, bool LOWER, int G, int PWN>
__global__ void __launch_bounds__(512) k_sweep_stage(const int* __restrict__ blev,
                                                     const int* __restrict__ bpos,
                                                     const int* __restrict__ lev_ptr_all,
                                                     const int* __restrict__ prs_all,
                                                     const int* __restrict__ pcol,
                                                     const VT* __restrict__ pval,
                                                     const VT* __restrict__ pdiag_all,
                                                     const int* __restrict__ gidx_all,
                                                     const double* __restrict__ gsrc,
                                                     const int* __restrict__ oidx_all,
                                                     double* __restrict__ odst,
                                                     double* __restrict__ xp_all,
                                                     const int* __restrict__ aidx,
                                                     const double* __restrict__ aval)
{
  __shared__ double ring[RING];
  __shared__ int s_lp[LPC + 2];
  const int b = blockIdx.x;
  const int lb = blev[b];
  const int nlev = blev[b + 1] - lb;
  const int* lev_ptr = lev_ptr_all + lb + 2 * b;   // nlev + 2 local offsets
  const int base = bpos[b];
  const int* prs = prs_all + base;
  const VT* pdiag = LOWER ? nullptr : pdiag_all + base;
  const int* gidx = gidx_all + base;
  const int* oidx = oidx_all ? oidx_all + base : nullptr;
  double* xp = xp_all + base;
  auto scatter = [&](int p, double v) {     // odst in equation order (+ the folded term)
    const int e = oidx[p];
    odst[e] = aidx ? v + aval[aidx[e]] : v;
  };
  const int t = threadIdx.x, nt = blockDim.x;
  const int grp = t / G, lane = t % G, ngr = nt / G;
  int lbase = 0;
  auto load_lp = [&](int l0) {
    for (int k = t; k < LPC + 2; k += nt) s_lp[k] = lev_ptr[min(l0 + k, nlev + 1)];
  };
  load_lp(0);
  __syncthreads();
  if (nlev <= 0) return;
  auto need_lp = [&](int L) {           // offsets of levels L and L + 1 staged
    if (L + 1 - lbase > LPC + 1) {
      __syncthreads();
      lbase = L - 1;
      load_lp(lbase);
      __syncthreads();
    }
  };
  int flushed = 0;                      // positions below are in the global copy
  // stage 1: row bounds and right-hand-side index of level [P_, P_ + nr_)
  auto stage1 = [&](int P_, int nr_, int& gi, int& qa, int& qb) {
    if (grp < nr_) {
      const int r = P_ + grp;
      gi = gidx[r];
      qa = prs[r];
      qb = prs[r + 1];
    }
  };
  // stage 2: right-hand side, diagonal, first PWN entries per lane
  auto stage2 = [&](int P_, int nr_, int gi, int qa, int qb, int* c, VT* v, double& in,
                    VT& dg) {
    if (grp < nr_) {
      in = gsrc[gi];
      if (!LOWER) dg = pdiag[P_ + grp];
#pragma unroll
      for (int j = 0; j < PWN; j++) {
        const int q = qa + lane + j * G;
        if (q < qb) { c[j] = pcol[q]; v[j] = pval[q]; }
        else { c[j] = -1; v[j] = (VT)0; }
      }
    }
  };
  int cN[PWN], cW[PWN];
  VT vN[PWN], vW[PWN];
  double inN = 0.0, inW = 0.0;
  VT dgN = (VT)1, dgW = (VT)1;
  int gi1 = 0, qa1 = 0, qb1 = 0, gi2 = 0, qa2 = 0, qb2 = 0, qaW = 0, qbW = 0;
#pragma unroll
  for (int j = 0; j < PWN; j++) { cN[j] = -1; vN[j] = (VT)0; cW[j] = -1; vW[j] = (VT)0; }
  // prologue: indices of levels 0 and 1, values of level 0
  int P0 = s_lp[0], nr0 = s_lp[1] - s_lp[0];
  int P1 = 0, nr1 = 0;
  stage1(P0, nr0, gi1, qa1, qb1);
  if (nlev > 1) {
    P1 = s_lp[1];
    nr1 = s_lp[2] - P1;
    stage1(P1, nr1, gi2, qa2, qb2);
  }
  stage2(P0, nr0, gi1, qa1, qb1, cN, vN, inN, dgN);
  for (int l = 0; l < nlev; l++) {
    // everything in flight was issued during the previous level
    __builtin_amdgcn_s_waitcnt(0x0F70);          // vmcnt(0)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < PWN; j++) { cW[j] = cN[j]; vW[j] = vN[j]; }
    inW = inN; dgW = dgN; qaW = qa1; qbW = qb1;
    gi1 = gi2; qa1 = qa2; qb1 = qb2;             // indices of level l + 1
    int P2 = 0, nr2 = 0;
    if (l + 2 < nlev) {
      need_lp(l + 2);
      P2 = s_lp[l + 2 - lbase];
      nr2 = s_lp[l + 3 - lbase] - P2;
    }
    __builtin_amdgcn_sched_barrier(0);
    if (l + 1 < nlev) stage2(P1, nr1, gi1, qa1, qb1, cN, vN, inN, dgN);
    if (l + 2 < nlev) stage1(P2, nr2, gi2, qa2, qb2);
    // level l: round 0 from the working registers, straight-line so that nothing issued
    // above is waited for; dependencies that left the ring, row tails beyond PWN entries
    // per lane and further rounds take the out-of-line path
    const int P = P0, nr = nr0;
    const bool wide = nr > WIDE;
    const int lim = wide ? INT_MAX : P + nr - RING;
    {
      const bool act = grp < nr;
      double s = 0.0;
      if (act) {
        bool old = false;
#pragma unroll
        for (int j = 0; j < PWN; j++) {
          const int cc = cW[j];
          if (cc >= 0) {
            old |= cc < lim;
            s += (double)vW[j] * ring[cc & (RING - 1)];
          }
        }
        if (old) {                           // dependencies that left the ring: global copy
          double xg[PWN];
#pragma unroll
          for (int j = 0; j < PWN; j++) {
            const int cc = cW[j];
            xg[j] = xp[(cc >= 0 && cc < lim) ? cc : 0];
          }
          s = 0.0;
#pragma unroll
          for (int j = 0; j < PWN; j++) {
            const int cc = cW[j];
            if (cc >= 0) {
              const double rv = ring[cc & (RING - 1)];
              s += (double)vW[j] * (cc < lim ? xg[j] : rv);
            }
          }
        }
        if (qbW > qaW + PWN * G)
          s += lane_sum<VT, G>(pval, pcol, ring, xp, qaW + lane + PWN * G, qbW, lim);
      }
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
      if (act && lane == 0) {
        const double xi = LOWER ? inW + s : (inW - s) / (double)dgW;
        ring[(P + grp) & (RING - 1)] = xi;
        if (wide) {
          xp[P + grp] = xi;
          if (oidx) scatter(P + grp, xi);
        }
      }
    }
    const int rounds = (nr + ngr - 1) / ngr;
    for (int k = 1; k < rounds; k++) {
      const int g = grp + k * ngr;
      const bool act = g < nr;
      double s = 0.0, in = 0.0;
      VT dg = (VT)1;
      if (act) {
        const int r = P + g;
        in = gsrc[gidx[r]];
        if (!LOWER) dg = pdiag[r];
        s = lane_sum<VT, G>(pval, pcol, ring, xp, prs[r] + lane, prs[r + 1], lim);
      }
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
      if (act && lane == 0) {
        const double xi = LOWER ? in + s : (in - s) / (double)dg;
        ring[(P + g) & (RING - 1)] = xi;
        if (wide) {
          xp[P + g] = xi;
          if (oidx) scatter(P + g, xi);
        }
      }
    }
    if (l + 1 >= nlev) break;
    const bool full = (P + nr) / FENCE != P / FENCE || wide || nr1 > WIDE;
    if (full) {
      if (!wide) {
        lds_barrier();
        for (int p = flushed + t; p < P + nr; p += nt) {
          const double v = ring[p & (RING - 1)];
          xp[p] = v;
          if (oidx) scatter(p, v);
        }
      }
      flushed = P + nr;
      __syncthreads();
    } else {
      lds_barrier();
    }
    P0 = P1; nr0 = nr1;
    P1 = P2; nr1 = nr2;
  }
  // the last level's results and whatever is still unflushed
  lds_barrier();
  if (nr0 <= WIDE)
    for (int p = flushed + t; p < P0 + nr0; p += nt) {
      const double v = ring[p & (RING - 1)];
      xp[p] = v;
      if (oidx) scatter(p, v);
    }
}

// ---------------------------------------------------------------------------------------
// block sweeps: one wave per block, both triangles in one launch

constexpr int BS_ROWS = 2047;      // rows per block (11-bit positions)
constexpr int BS_LROWS = 32;       // rows per level (5-bit row numbers)
constexpr int BS_LPAIRS = 64;      // entry pairs per level (one per lane)
constexpr int BS_PAD = 256;        // padding after every entry stream (clamped loads)
constexpr unsigned BS_EM = (1u << 21) - 1;
constexpr int BS_DEPTH = 4;        // levels whose loads are in flight ahead of the level
constexpr int BS_GATHER = 8;       // rows per lane whose loads are issued together

template <class VT> struct VPair;
template <> struct VPair<float> { using T = float2; using Q = float4; };
template <> struct VPair<double> { using T = double2; using Q = double4; };

// One step of an inclusive scan over lanes segmented by key (equal keys are contiguous,
// keys >= 1): every lane adds the value DPP control C brings from another lane when that
// lane has the same key and the step applies to the lane's row (ok).  A lane the control
// leaves without a source reads zeros (key 0: no match).
template <int C>
__device__ __forceinline__ void seg_step(double& v, int key, bool ok)
{
  const int ks = __builtin_amdgcn_mov_dpp(key, C, 0xF, 0xF, true);
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), C, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), C, 0xF, 0xF, true);
  const double vs = __hiloint2double(hi, lo);
  v = (ks == key && ok) ? v + vs : v;
}

// Sum of v over each run of equal keys, complete at the run's last lane: shifts by 1, 2, 4
// and 8 inside each row of 16 lanes, then lane 15 of rows 0 and 2 into rows 1 and 3, then
// lane 31 into rows 2 and 3 (each row already holds its own part when it is passed on).
__device__ __forceinline__ double seg_sum(double v, int key, bool rows13, bool rows23)
{
  seg_step<0x111>(v, key, true);     // row_shr:1
  seg_step<0x112>(v, key, true);     // row_shr:2
  seg_step<0x114>(v, key, true);     // row_shr:4
  seg_step<0x118>(v, key, true);     // row_shr:8
  seg_step<0x142>(v, key, rows13);   // row_bcast:15
  seg_step<0x143>(v, key, rows23);   // row_bcast:31
  return v;
}

// The levels of one triangle of one block.  A level lists nodes (units of 1-3 rows); its
// entries are the couplings of those rows to rows outside their node, a row's entries
// padded to an even count (at least one pair), lane L taking pair L.  Loaded BS_DEPTH
// levels ahead into a register ring: the value pair, the packed pair of column | row-in-
// level words, the level's table entries (kept uniform) and, for lane j, node j's record
// and couplings.  A level: every lane multiplies its pair by the solved values of the
// columns, a segmented scan sums each row at its last lane, which leaves the sum in LDS;
// then lane j solves node j's rows in order with the couplings inside the node: the lower
// sweep x_r += sum_r + sum of l_rs x_s (factors stored negated), the upper sweep (rows in
// reverse) x_r = (x_r - sum_r - sum of u_rs x_s) / u_rr.  One wave works the block, so a
// level ends without a workgroup barrier.
template <class VT, bool LOWER>
__device__ __forceinline__ void block_levels(int nlev, const unsigned* tab,
                                             const VT* __restrict__ vals,
                                             const unsigned short* __restrict__ meta,
                                             const unsigned* __restrict__ unit,
                                             const VT* __restrict__ uval,
                                             const double4* __restrict__ rdg, double* x,
                                             double* acc, double* sink, int lane, int m)
{
  using V2 = typename VPair<VT>::T;
  using V4 = typename VPair<VT>::Q;
  constexpr int D = BS_DEPTH;
  const V2* vp = reinterpret_cast<const V2*>(vals);           // entry bases are even
  const unsigned* mp = reinterpret_cast<const unsigned*>(meta);
  const V4* up4 = reinterpret_cast<const V4*>(uval);          // 4 couplings per node
  V2 vv[D];
  V4 ua[D];
  double4 ub[D];
  unsigned cm[D], ui[D], ta[D], tb[D];
  const bool rows13 = (lane >> 4) & 1, rows23 = lane >= 32;
#pragma unroll
  for (int k = 0; k < D; k++) {                            // the first D levels are empty
    ta[k] = 0; tb[k] = 0; cm[k] = 0; ui[k] = 1u << 11;
    vv[k].x = (VT)0; vv[k].y = (VT)0;
  }
  // stages run for levels 0, 1, 2, ...; past the last level the doubled sentinel makes an
  // empty level with the same loads.  The table entry after the stage's level is read one
  // stage ahead, so no stage waits for it.
  unsigned t_lo = __builtin_amdgcn_readfirstlane(tab[0]);
  unsigned t_hi_v = tab[min(1, nlev + 1)];
  int ls = 0;
  auto stage = [&](int k) {
    const unsigned t0 = t_lo, t1 = __builtin_amdgcn_readfirstlane(t_hi_v);
    t_hi_v = tab[min(ls + 2, nlev + 1)];
    t_lo = t1;
    ls++;
    ta[k] = t0; tb[k] = t1;
    const int e0 = (int)(t0 & BS_EM), np = ((int)(t1 & BS_EM) - e0) >> 1;
    const int q = (e0 >> 1) + min(lane, max(np - 1, 0));
    vv[k] = vp[q];
    cm[k] = mp[q];
    const int U0 = (int)(t0 >> 21), nu = (int)(t1 >> 21) - U0;
    const int u = U0 + min(lane, max(nu - 1, 0));
    ui[k] = unit[u];
    ua[k] = up4[u];
    if (!LOWER) ub[k] = rdg[u];
  };
  auto level = [&](int k) {
    const unsigned t0 = ta[k], t1 = tb[k];
    const int np = ((int)(t1 & BS_EM) - (int)(t0 & BS_EM)) >> 1;
    const int nu = (int)(t1 >> 21) - (int)(t0 >> 21);
    const unsigned ca = cm[k] & 0xffffu, cb = cm[k] >> 16;   // one row: the same row bits
    double v = (double)vv[k].x * x[ca & 2047u] + (double)vv[k].y * x[cb & 2047u];
    const bool on = lane < np;
    const int key = on ? 1 + (int)(ca >> 11) : 64;
    v = on ? v : 0.0;
    v = seg_sum(v, key, rows13, rows23);
    const int kn = __builtin_amdgcn_mov_dpp(key, 0x130, 0xF, 0xF, true);   // wave_shl:1
    if (on && kn != key) acc[key - 1] = v;
    wave_sync();
    if (lane < nu) {
      // branch-free: rows past the node's last read valid slots, take neutral operands and
      // store to the sink
      const unsigned inf = ui[k];
      const int p = (int)(inf & 2047u), nr = (int)((inf >> 11) & 3u), s0 = (int)(inf >> 13);
      const bool h1 = nr > 1, h2 = nr > 2;
      const int r1 = h1 ? 1 : 0, r2 = h2 ? 2 : r1;
      const double a0 = acc[s0], a1 = acc[s0 + r1], a2 = acc[s0 + r2];
      const double x0 = x[p], x1 = x[p + r1], x2 = x[p + r2];
      double y0, y1, y2;
      if (LOWER) {
        // couplings l10, l20, l21 (zero where absent)
        y0 = x0 + a0;
        y1 = x1 + a1 + (double)ua[k].x * y0;
        y2 = x2 + a2 + (double)ua[k].y * y0 + (double)ua[k].z * y1;
      } else {
        // couplings u01, u02, u12 (zero where absent), reciprocal diagonals (1 where
        // absent); last row first
        y2 = (h2 ? x2 - a2 : 0.0) * ub[k].z;
        y1 = (h1 ? x1 - a1 - (double)ua[k].z * y2 : 0.0) * ub[k].y;
        y0 = (x0 - a0 - (double)ua[k].x * y1 - (double)ua[k].y * y2) * ub[k].x;
      }
      x[p] = y0;
      *(h1 ? x + p + 1 : sink) = y1;
      *(h2 ? x + p + 2 : sink) = y2;
    }
    wave_sync();
  };
  // no branch around a stage, so that every level issues the same loads and the wait
  // before a level's first use of its registers leaves the later levels' loads in flight;
  // the first D (empty) levels issue the loads of levels 0 .. D - 1, so that every load
  // of the ring is issued inside the loop, into the same registers
  for (int l0 = -D; l0 < nlev; l0 += D) {
#pragma unroll
    for (int k = 0; k < D; k++) {
      level(k);
      stage(k);
    }
  }
}

// One block per one-wave workgroup: the block's right-hand side is gathered into LDS (in
// local positions), swept down and up in place, and written back in equation order (with
// aidx, plus aval[aidx[e]]: the coarse prolongation folded into the write).  Blocks hold
// disjoint rows, so b and x may be the same vector.
template <class VT>
__global__ void __launch_bounds__(64) k_block_sweep(const int4* __restrict__ info,
                                                    const int* __restrict__ gidx,
                                                    const unsigned* __restrict__ levt,
                                                    const VT* __restrict__ vlo,
                                                    const unsigned short* __restrict__ mlo,
                                                    const unsigned* __restrict__ ulo,
                                                    const VT* __restrict__ wlo,
                                                    const VT* __restrict__ vup,
                                                    const unsigned short* __restrict__ mup,
                                                    const unsigned* __restrict__ uup,
                                                    const VT* __restrict__ wup,
                                                    const double4* __restrict__ rdg,
                                                    const double* rhs, double* out,
                                                    const int* __restrict__ aidx,
                                                    const double* __restrict__ aval)
{
  constexpr int G = BS_GATHER;
  extern __shared__ double smem[];
  const int4 i0 = info[2 * blockIdx.x], i1 = info[2 * blockIdx.x + 1];
  const int rb = i0.x, m = i0.y, lb = i0.z, nl = i0.w, nu = i1.x, ub = i1.y;
  const int lane = threadIdx.x;
  double* x = smem;                                           // [m] local positions
  double* acc = smem + m;                                     // [BS_LROWS] row sums
  double* sink = acc + BS_LROWS;                              // [2] stores of absent rows
  unsigned* tl = reinterpret_cast<unsigned*>(sink + 2);       // [nl + 2]
  unsigned* tu = tl + nl + 2;                                 // [nu + 2]
  // gathers: every load of a chunk issued before any is waited for
  for (int k0 = 0; k0 < m; k0 += 64 * G) {
    int g[G];
    double v[G];
#pragma unroll
    for (int j = 0; j < G; j++) g[j] = gidx[rb + min(k0 + lane + 64 * j, m - 1)];
#pragma unroll
    for (int j = 0; j < G; j++) v[j] = rhs[g[j]];
#pragma unroll
    for (int j = 0; j < G; j++)
      if (k0 + lane + 64 * j < m) x[k0 + lane + 64 * j] = v[j];
  }
  const int nt = nl + nu + 4;                                 // both tables, contiguous
  for (int k0 = 0; k0 < nt; k0 += 64 * G) {
    unsigned t[G];
#pragma unroll
    for (int j = 0; j < G; j++) t[j] = levt[lb + min(k0 + lane + 64 * j, nt - 1)];
#pragma unroll
    for (int j = 0; j < G; j++)
      if (k0 + lane + 64 * j < nt) tl[k0 + lane + 64 * j] = t[j];
  }
  wave_sync();
  block_levels<VT, true>(nl, tl, vlo + i1.z, mlo + i1.z, ulo + ub, wlo + 4 * (long long)ub,
                         nullptr, x, acc, sink, lane, m);
  block_levels<VT, false>(nu, tu, vup + i1.w, mup + i1.w, uup + ub, wup + 4 * (long long)ub,
                          rdg + ub, x, acc, sink, lane, m);
  wave_sync();
  for (int k0 = 0; k0 < m; k0 += 64 * G) {
    int e[G];
    double a[G];
#pragma unroll
    for (int j = 0; j < G; j++) e[j] = gidx[rb + min(k0 + lane + 64 * j, m - 1)];
    if (aidx) {
      int ai[G];
#pragma unroll
      for (int j = 0; j < G; j++) ai[j] = aidx[e[j]];
#pragma unroll
      for (int j = 0; j < G; j++) a[j] = aval[ai[j]];
    }
#pragma unroll
    for (int j = 0; j < G; j++) {
      const int k = k0 + lane + 64 * j;
      if (k < m) out[e[j]] = aidx ? x[k] + a[j] : x[k];
    }
  }
}

// ---------------------------------------------------------------------------------------
// host set-up

// Level sets of the lower or upper dependency graph of rows [a, e), entries with a column
// outside [a, e) or (with `seg`) in another segment ignored.
static void levels(int a, int e, const std::vector<int>& rp, const std::vector<int>& cl, bool lower,
                   const std::vector<int>& seg, std::vector<int>& lev, std::vector<int>& ptr,
                   std::vector<int>& rows)
{
  int nlev = 0;
  const bool sg = !seg.empty();
  if (lower) {
    for (int i = a; i < e; i++) {
      int l = 0;
      for (int q = rp[i] + 1; q < rp[i + 1]; q++)
        if (cl[q] < i && cl[q] >= a && (!sg || seg[cl[q]] == seg[i])) l = std::max(l, lev[cl[q]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  } else {
    for (int i = e - 1; i >= a; i--) {
      int l = 0;
      for (int q = rp[i] + 1; q < rp[i + 1]; q++)
        if (cl[q] > i && cl[q] < e && (!sg || seg[cl[q]] == seg[i])) l = std::max(l, lev[cl[q]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  }
  ptr.assign(nlev + 1, 0);
  for (int i = a; i < e; i++) ptr[lev[i] + 1]++;
  for (int l = 0; l < nlev; l++) ptr[l + 1] += ptr[l];
  rows.assign(e - a, 0);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  if (lower) { for (int i = a; i < e; i++) rows[fill[lev[i]]++] = i; }
  else { for (int i = e - 1; i >= a; i--) rows[fill[lev[i]]++] = i; }
}

// Packed layout of one sweep over all blocks.
static void build_sweep(SweepSet& S, const Ilu& I, bool lower, std::vector<int>& order,
                        hipStream_t s, int cap)
{
  const int n = I.n, nb = I.nblk;
  const std::vector<int>& rp = I.h_rp;
  const std::vector<int>& cl = I.h_cols;
  std::vector<int> lev(n, 0);
  std::vector<int> blev(nb + 1, 0), bpos(nb + 1, 0), levp, prow;
  levp.reserve(n / 8);
  prow.reserve(n);
  int maxnr = 0, nlev_tot = 0;
  for (int b = 0; b < nb; b++) {
    const int a = I.bstart[b], e = I.bstart[b + 1];
    std::vector<int> ptr, rows;
    levels(a, e, rp, cl, lower, I.seg, lev, ptr, rows);
    if (cap > 0) {                                 // levels of more than cap rows split
      std::vector<int> cp(1, 0);
      for (size_t l = 0; l + 1 < ptr.size(); l++)
        for (int r = ptr[l]; r < ptr[l + 1]; r += cap) cp.push_back(std::min(r + cap, ptr[l + 1]));
      ptr.swap(cp);
    }
    const int nl = (int)ptr.size() - 1;
    blev[b + 1] = blev[b] + nl;
    bpos[b + 1] = bpos[b] + (e - a);
    for (int l = 0; l <= nl; l++) levp.push_back(ptr[l]);
    levp.push_back(ptr[nl]);                       // sentinel: lev_ptr[l+2] for the last level
    for (int l = 0; l < nl; l++) maxnr = std::max(maxnr, ptr[l + 1] - ptr[l]);
    prow.insert(prow.end(), rows.begin(), rows.end());
    nlev_tot = std::max(nlev_tot, nl);
  }
  // positions: local to each block
  std::vector<int> pos(n);
  for (int b = 0; b < nb; b++)
    for (int r = bpos[b]; r < bpos[b + 1]; r++) pos[prow[r]] = r - bpos[b];
  // entries of row i kept by the sweep: the strict lower / upper part inside i's strip
  std::vector<int> blk(n);
  for (int b = 0; b < nb; b++)
    for (int i = I.bstart[b]; i < I.bstart[b + 1]; i++) blk[i] = b;
  const bool sg = !I.seg.empty();
  auto keep = [&](int i, int c) {
    return (lower ? c < i : c > i) && c >= I.bstart[blk[i]] && c < I.bstart[blk[i] + 1] &&
           (!sg || I.seg[c] == I.seg[i]);
  };
  std::vector<int> prs(n + 1, 0);
  for (int r = 0; r < n; r++) {
    const int i = prow[r];
    int k = 0;
    for (int q = rp[i] + 1; q < rp[i + 1]; q++) k += keep(i, cl[q]);
    prs[r + 1] = prs[r] + k;
  }
  const long long ne = prs[n];
  std::vector<int> pcol(ne), psrc(ne);
  parallel_for(n, [&](long long a, long long b) {
    for (long long r = a; r < b; r++) {
      const int i = prow[r];
      int j = prs[r];
      for (int q = rp[i] + 1; q < rp[i + 1]; q++)
        if (keep(i, cl[q])) { pcol[j] = pos[cl[q]]; psrc[j] = q; j++; }
    }
  });
  S.nblk = nb;
  S.nlev_max = nlev_tot;
  S.max_nr = maxnr;
  S.ne = ne;
  S.nlev_sum = blev[nb];
  if (getenv("RMX_DEBUG"))
    fprintf(stderr, "[rmx] %s sweep, %d strip(s) x %d segment(s): %d levels (longest strip), "
            "%.1f rows/level, max %d, %lld entries\n", lower ? "lower" : "upper", nb, I.nseg,
            nlev_tot, (double)n / std::max(1, blev[nb]), maxnr, ne);
  S.blev.upload(blev, s); S.bpos.upload(bpos, s); S.lev_ptr.upload(levp, s);
  S.prs.upload(prs, s); S.prow.upload(prow, s); S.pcol.upload(pcol, s); S.psrc.upload(psrc, s);
  S.xp.alloc(n);
  order = prow;
}

// Block sweep layout.  A block is the rows of one strip in one segment; a unit is a run of
// rows of one node (group), at most 3 (every row its own unit without `group`).  Unit
// levels: a unit's lower level is one more than the highest of the units its rows couple to
// below them in the block (upper: above); units of a level are independent.  Local
// positions list the units in lower level order, a unit's rows together.  Levels are split
// into parts of at most BS_LROWS rows and BS_LPAIRS entry pairs.  Entries keep the strip
// sweeps' rule (the strict triangle inside the row's block), minus the couplings inside a
// unit, which go to the unit's record.  Off (the strip sweeps run) when a block has more than
// BS_ROWS rows or its entries do not fit 21 bits.
static void build_blocks(Ilu& I, const std::vector<int>* group, hipStream_t s)
{
  BlockSweep& B = I.bs;
  B.on = false;
  const int n = I.n, nstrip = I.nblk, nseg = std::max(1, I.nseg);
  const std::vector<int>& rp = I.h_rp;
  const std::vector<int>& cl = I.h_cols;
  const bool sg = !I.seg.empty();
  auto segof = [&](int i) { return sg ? I.seg[i] : 0; };
  if (nseg == 1)
    for (int b = 0; b < nstrip; b++)
      if (I.bstart[b + 1] - I.bstart[b] > BS_ROWS) return;
  // tighter level caps (tests of the level split): RMX_BLOCK_LEVEL_ROWS, RMX_BLOCK_LEVEL_PAIRS
  int cap_rows = BS_LROWS, cap_pairs = BS_LPAIRS;
  if (const char* e = getenv("RMX_BLOCK_LEVEL_ROWS")) cap_rows = std::max(3, std::min(BS_LROWS, atoi(e)));
  if (const char* e = getenv("RMX_BLOCK_LEVEL_PAIRS")) cap_pairs = std::max(1, std::min(BS_LPAIRS, atoi(e)));
  // CSR position of entry (r, c), -1 if the pattern has none
  auto find = [&](int r, int c) {
    for (int q = rp[r]; q < rp[r + 1]; q++)
      if (cl[q] == c) return q;
    return -1;
  };
  struct Part {
    std::vector<int> rows;               // equation at every local position
    std::vector<unsigned> tab[2];        // level tables (lower, upper)
    std::vector<unsigned short> meta[2];
    std::vector<int> src[2];
    std::vector<unsigned> unit[2];       // unit records
    std::vector<int> usrc[2];            // unit coupling sources (4 per unit)
    std::vector<int> dsrc;               // unit diagonal sources (4 per unit)
    bool fits = true;
  };
  std::vector<Part> parts((size_t)nstrip * nseg);
  parallel_for(nstrip, [&](long long b0, long long b1) {
    for (long long b = b0; b < b1; b++) {
      const int a = I.bstart[b], e = I.bstart[b + 1];
      // units of the strip (runs of one node's rows in one segment, at most 3)
      std::vector<int> ustart, uof(e - a);
      for (int i = a; i < e; i++) {
        const bool same = i > a && group && (*group)[i] == (*group)[i - 1] &&
                          segof(i) == segof(i - 1) && i - ustart.back() < 3;
        if (!same) ustart.push_back(i);
        uof[i - a] = (int)ustart.size() - 1;
      }
      auto inblock = [&](int i, int c) { return c >= a && c < e && segof(c) == segof(i); };
      // entry pairs of a row (at least one) when its unit starts at u0 and ends before u1
      auto row_pairs = [&](int i, int u0, int u1, bool lower) {
        int w = 0;
        for (int q = rp[i] + 1; q < rp[i + 1]; q++) {
          const int c = cl[q];
          w += (lower ? c < i : c > i) && inblock(i, c) && (c < u0 || c >= u1);
        }
        return std::max(1, (w + 1) / 2);
      };
      // a unit whose rows need more pairs than a level holds becomes single-row units (a
      // row has at most 127 entries, 64 pairs)
      {
        std::vector<int> us2;
        for (size_t U = 0; U < ustart.size(); U++) {
          const int u0 = ustart[U], u1 = U + 1 < ustart.size() ? ustart[U + 1] : e;
          int pl = 0, pu = 0;
          for (int i = u0; i < u1; i++) { pl += row_pairs(i, u0, u1, true); pu += row_pairs(i, u0, u1, false); }
          if (std::max(pl, pu) <= BS_LPAIRS) us2.push_back(u0);
          else for (int i = u0; i < u1; i++) us2.push_back(i);
        }
        ustart.swap(us2);
        for (size_t U = 0; U < ustart.size(); U++) {
          const int u1 = U + 1 < ustart.size() ? ustart[U + 1] : e;
          for (int i = ustart[U]; i < u1; i++) uof[i - a] = (int)U;
        }
      }
      const int nu = (int)ustart.size();
      ustart.push_back(e);
      // unit levels, lower and upper
      std::vector<int> ulev[2] = { std::vector<int>(nu, 0), std::vector<int>(nu, 0) };
      for (int U = 0; U < nu; U++)
        for (int i = ustart[U]; i < ustart[U + 1]; i++)
          for (int q = rp[i] + 1; q < rp[i + 1]; q++) {
            const int c = cl[q];
            if (c < i && inblock(i, c) && uof[c - a] != U)
              ulev[0][U] = std::max(ulev[0][U], ulev[0][uof[c - a]] + 1);
          }
      for (int U = nu - 1; U >= 0; U--)
        for (int i = ustart[U]; i < ustart[U + 1]; i++)
          for (int q = rp[i] + 1; q < rp[i + 1]; q++) {
            const int c = cl[q];
            if (c > i && inblock(i, c) && uof[c - a] != U)
              ulev[1][U] = std::max(ulev[1][U], ulev[1][uof[c - a]] + 1);
          }
      // units of every segment
      std::vector<std::vector<int>> segu(nseg);
      for (int U = 0; U < nu; U++) segu[segof(ustart[U])].push_back(U);
      for (int g = 0; g < nseg; g++) {
        Part& P = parts[(size_t)b * nseg + g];
        std::vector<int>& us = segu[g];
        if (us.empty()) continue;
        int m = 0;
        for (int U : us) m += ustart[U + 1] - ustart[U];
        if (m > BS_ROWS) { P.fits = false; continue; }
        // lower order: by lower level, then unit index; local positions follow it
        std::vector<int> ord[2] = { us, us };
        for (int t = 0; t < 2; t++)
          std::stable_sort(ord[t].begin(), ord[t].end(),
                           [&](int x, int y) { return ulev[t][x] < ulev[t][y]; });
        std::vector<int> upos(nu, -1);
        for (int U : ord[0]) {
          upos[U] = (int)P.rows.size();
          for (int i = ustart[U]; i < ustart[U + 1]; i++) P.rows.push_back(i);
        }
        auto pos = [&](int c) { return upos[uof[c - a]] + (c - ustart[uof[c - a]]); };
        for (int t = 0; t < 2; t++) {
          const bool lower = t == 0;
          std::vector<unsigned>& tab = P.tab[t];
          std::vector<unsigned short>& meta = P.meta[t];
          std::vector<int>& src = P.src[t];
          int cur = -1, nrow = 0, npair = 0;
          for (int k = 0; k < (int)ord[t].size(); k++) {
            const int U = ord[t][k], r0 = ustart[U], nr = ustart[U + 1] - r0;
            int pairs = 0;                      // the unit's entry pairs (a row at least one)
            for (int i = r0; i < r0 + nr; i++) pairs += row_pairs(i, r0, r0 + nr, lower);
            if (ulev[t][U] != cur || nrow + nr > cap_rows || npair + pairs > cap_pairs) {
              tab.push_back(((unsigned)k << 21) | (unsigned)meta.size());
              cur = ulev[t][U]; nrow = 0; npair = 0;
            }
            for (int j = 0; j < nr; j++) {
              const int i = r0 + j;
              const size_t start = meta.size();
              for (int q = rp[i] + 1; q < rp[i + 1]; q++) {
                const int c = cl[q];
                if ((lower ? c < i : c > i) && inblock(i, c) && uof[c - a] != U) {
                  meta.push_back((unsigned short)(pos(c) | ((nrow + j) << 11)));
                  src.push_back(q);
                }
              }
              while (meta.size() - start < 2 || ((meta.size() - start) & 1)) {   // zero entries
                meta.push_back((unsigned short)(((nrow + j) << 11)));
                src.push_back(-1);
              }
            }
            P.unit[t].push_back((unsigned)upos[U] | ((unsigned)nr << 11) | ((unsigned)nrow << 13));
            const int r1 = r0 + std::min(1, nr - 1), r2 = r0 + std::min(2, nr - 1);
            if (lower) {
              const int l10 = nr > 1 ? find(r1, r0) : -1, l20 = nr > 2 ? find(r2, r0) : -1;
              const int l21 = nr > 2 ? find(r2, r1) : -1;
              for (int v : { l10, l20, l21, -1 }) P.usrc[t].push_back(v);
            } else {
              const int u01 = nr > 1 ? find(r0, r1) : -1, u02 = nr > 2 ? find(r0, r2) : -1;
              const int u12 = nr > 2 ? find(r1, r2) : -1;
              for (int v : { u01, u02, u12, -1 }) P.usrc[t].push_back(v);
              for (int v : { rp[r0], nr > 1 ? rp[r1] : -1, nr > 2 ? rp[r2] : -1, -1 })
                P.dsrc.push_back(v);
            }
            nrow += nr;
            npair += pairs;
          }
          tab.push_back(((unsigned)ord[t].size() << 21) | (unsigned)meta.size());
          tab.push_back(tab.back());         // read by the prefetch past the last level
          for (int z = 0; z < 2; z++) {      // a zero pair, loaded by the levels past the last
            meta.push_back(0);
            src.push_back(-1);
          }
          if (meta.size() > BS_EM) P.fits = false;
        }
      }
    }
  });
  long long nb = 0, ent[2] = {0, 0}, nlev = 0, nunit = 0;
  size_t lds = 0;
  for (const Part& P : parts) {
    if (!P.fits) return;
    if (P.rows.empty()) continue;
    nb++;
    for (int t = 0; t < 2; t++) ent[t] += P.meta[t].size();
    nlev += P.tab[0].size() + P.tab[1].size();
    nunit += P.unit[0].size();
    const size_t m = P.rows.size();
    lds = std::max(lds, 8 * (m + BS_LROWS + 2) + 4 * (P.tab[0].size() + P.tab[1].size()));
  }
  if (nb == 0 || ent[0] + BS_PAD > INT_MAX || ent[1] + BS_PAD > INT_MAX) return;
  std::vector<int4> info(2 * nb);
  std::vector<int> gidx(n);
  std::vector<int> src[2], usrc[2];
  std::vector<unsigned short> meta[2];
  std::vector<unsigned> unit[2];
  for (int t = 0; t < 2; t++) {
    src[t].assign(ent[t] + BS_PAD, -1);
    meta[t].assign(ent[t] + BS_PAD, 0);
    unit[t].assign(nunit + 64, 1u << 11);
    usrc[t].assign((size_t)4 * (nunit + 64), -1);
  }
  std::vector<int> dsrc((size_t)4 * (nunit + 64), -1);
  std::vector<unsigned> lev(nlev);
  long long k = 0, rb = 0, lb = 0, eo[2] = {0, 0}, ub = 0;
  for (const Part& P : parts) {
    if (P.rows.empty()) continue;
    const int m = (int)P.rows.size();
    const int nl = (int)P.tab[0].size() - 2, nu = (int)P.tab[1].size() - 2;
    info[2 * k] = int4{ (int)rb, m, (int)lb, nl };
    info[2 * k + 1] = int4{ nu, (int)ub, (int)eo[0], (int)eo[1] };
    std::copy(P.tab[0].begin(), P.tab[0].end(), lev.begin() + lb);
    std::copy(P.tab[1].begin(), P.tab[1].end(), lev.begin() + lb + nl + 2);
    std::copy(P.rows.begin(), P.rows.end(), gidx.begin() + rb);
    for (int t = 0; t < 2; t++) {
      std::copy(P.meta[t].begin(), P.meta[t].end(), meta[t].begin() + eo[t]);
      std::copy(P.src[t].begin(), P.src[t].end(), src[t].begin() + eo[t]);
      std::copy(P.unit[t].begin(), P.unit[t].end(), unit[t].begin() + ub);
      std::copy(P.usrc[t].begin(), P.usrc[t].end(), usrc[t].begin() + 4 * ub);
      eo[t] += P.meta[t].size();
    }
    std::copy(P.dsrc.begin(), P.dsrc.end(), dsrc.begin() + 4 * ub);
    k++; rb += m; lb += nl + nu + 4; ub += P.unit[0].size();
  }
  B.nb = (int)nb; B.lds = (lds + 15) & ~(size_t)15;
  B.ne_lo = std::count_if(src[0].begin(), src[0].end(), [](int q) { return q >= 0; });
  B.ne_up = std::count_if(src[1].begin(), src[1].end(), [](int q) { return q >= 0; });
  B.nlev = nlev; B.nunit = nunit;
  B.info.upload(info, s); B.gidx.upload(gidx, s); B.lev.upload(lev, s);
  B.mlo.upload(meta[0], s); B.mup.upload(meta[1], s);
  B.slo.upload(src[0], s); B.sup.upload(src[1], s);
  B.ulo.upload(unit[0], s); B.uup.upload(unit[1], s);
  B.sulo.upload(usrc[0], s); B.suup.upload(usrc[1], s); B.sdg.upload(dsrc, s);
  HIPCHECK(hipStreamSynchronize(s));
  B.on = true;
  if (getenv("RMX_DEBUG"))
    fprintf(stderr, "[rmx] block sweeps: %lld blocks, %lld nodes, %lld + %lld entries (%lld + "
            "%lld padded), %.1f levels per block, %zu bytes LDS\n", nb, nunit, B.ne_lo, B.ne_up,
            ent[0], ent[1], (double)nlev / nb - 4.0, B.lds);
}

// Segment of every row: each strip [bstart[b], bstart[b+1]) is cut into nseg parts of
// about equal row count along the principal axis of its nodes' coordinates (across the
// channel for a strip of cross-sections); a node's rows share a segment.
std::vector<int> cut_segments(const std::vector<int>& bstart, const std::vector<int>& group,
                              const HostModel& hm, int nseg)
{
  const int nblk = (int)bstart.size() - 1;
  std::vector<int> seg(bstart[nblk], 0);
  parallel_for(nblk, [&](long long b0, long long b1) {
    std::vector<std::pair<double, int>> proj;
    for (long long b = b0; b < b1; b++) {
      const int a = bstart[b], e = bstart[b + 1];
      if (e <= a) continue;
      double mx = 0, my = 0;
      for (int i = a; i < e; i++) { mx += hm.x[group[i]]; my += hm.y[group[i]]; }
      mx /= (e - a); my /= (e - a);
      double sxx = 0, sxy = 0, syy = 0;
      for (int i = a; i < e; i++) {
        const double dx = hm.x[group[i]] - mx, dy = hm.y[group[i]] - my;
        sxx += dx * dx; sxy += dx * dy; syy += dy * dy;
      }
      // eigenvector of the larger eigenvalue of [[sxx, sxy], [sxy, syy]]
      const double th = 0.5 * std::atan2(2.0 * sxy, sxx - syy);
      const double ex = std::cos(th), ey = std::sin(th);
      proj.clear();
      for (int i = a; i < e; i++)
        proj.emplace_back((hm.x[group[i]] - mx) * ex + (hm.y[group[i]] - my) * ey, i);
      std::sort(proj.begin(), proj.end());
      // equal row counts, cuts moved to the next node so a node's rows share a segment
      const int m = e - a;
      int k = 0;
      for (int t = 0; t < m; t++) {
        if (k + 1 < nseg && t >= (long long)m * (k + 1) / nseg && t > 0 &&
            group[proj[t].second] != group[proj[t - 1].second] && proj[t].first != proj[t - 1].first)
          k++;
        seg[proj[t].second] = k;
      }
    }
  });
  return seg;
}

// Strips: contiguous row ranges of about equal size.  With `group` (the node of every
// equation) a cut moves forward to the next row of another node, so the U, V, H equations
// of one node stay in one strip.  level_cap > 0 splits the sweeps' levels into parts of at
// most that many rows.
void ilu_setup(Ilu& I, int n, const std::vector<int>& rp, const std::vector<int>& cl, int nblk,
               const std::vector<int>* group, hipStream_t s, const HostModel* hm, int nseg,
               int level_cap)
{
  I.ready = I.fac64 = I.fac32 = false;
  I.n = n;
  I.nblk = std::max(1, std::min(nblk, n));
  I.bstart.assign(I.nblk + 1, 0);
  for (int b = 0; b <= I.nblk; b++) {
    int cut = (int)((long long)n * b / I.nblk);
    if (group && b > 0) {
      cut = std::max(cut, I.bstart[b - 1]);
      while (cut > 0 && cut < n && (*group)[cut] == (*group)[cut - 1]) cut++;
    }
    I.bstart[b] = cut;
  }
  I.bstart[I.nblk] = n;
  // segments: each strip is cut across into nseg parts of equal row count along the
  // principal axis of its nodes' coordinates (across the channel for a strip of
  // cross-sections); entries coupling two segments are left out of the sweeps as well
  I.seg.clear();
  I.nseg = 1;
  if (group && hm && nseg > 1) {
    I.nseg = nseg;
    I.seg = cut_segments(I.bstart, *group, *hm, nseg);
  }
  I.h_rp = rp;
  I.h_cols = cl;
  I.nnz = I.h_rp[n];
  // factorisation schedule: lower levels of the whole pattern
  std::vector<int> lev(n, 0), fptr, frows;
  levels(0, n, I.h_rp, I.h_cols, true, std::vector<int>(), lev, fptr, frows);
  I.flev = fptr;
  I.flev_rows.upload(frows, s);
  // per-row sorted columns for the factorisation lookups
  std::vector<int> scol(I.nnz);
  std::vector<unsigned char> spos(I.nnz);
  for (int i = 0; i < n; i++)
    if (I.h_rp[i + 1] - I.h_rp[i] > FW) fail(RMX_ERR_RANGE, "ILU(0): row wider than 128 columns");
  parallel_for(n, [&](long long a, long long b) {
    std::vector<std::pair<int, int>> tmp;
    for (long long i = a; i < b; i++) {
      tmp.clear();
      for (int q = I.h_rp[i]; q < I.h_rp[i + 1]; q++) tmp.emplace_back(I.h_cols[q], q - I.h_rp[i]);
      std::sort(tmp.begin(), tmp.end());
      for (size_t t = 0; t < tmp.size(); t++) {
        scol[I.h_rp[i] + t] = tmp[t].first;
        spos[I.h_rp[i] + t] = (unsigned char)tmp[t].second;
      }
    }
  });
  I.rp.upload(I.h_rp, s); I.cols.upload(I.h_cols, s);
  I.srt_cols.upload(scol, s); I.srt_pos.upload(spos, s);
  // factorisation metadata: per level position {row, start, width, lower entries}; per
  // sorted lower entry the start and length (after the diagonal) of its pivot row
  {
    std::vector<int4> desc(n);
    std::vector<int2> piv(I.nnz, int2{0, 0});
    parallel_for(n, [&](long long a, long long e) {
      for (long long w = a; w < e; w++) {
        const int i = frows[w];
        const int p0 = I.h_rp[i], wi = I.h_rp[i + 1] - p0;
        int nk = 0;
        while (nk < wi && scol[p0 + nk] < i) {
          const int k = scol[p0 + nk];
          piv[p0 + nk] = int2{ I.h_rp[k], I.h_rp[k + 1] - I.h_rp[k] - 1 };
          nk++;
        }
        desc[w] = int4{ i, p0, wi, nk };
      }
    });
    // plan offsets: every row's concatenated pivot-row entries
    std::vector<long long> pbase(n + 1, 0);
    for (int w = 0; w < n; w++) {
      long long tot = 0;
      for (int t = 0; t < desc[w].w; t++) tot += piv[desc[w].y + t].y;
      pbase[w + 1] = pbase[w] + tot;
    }
    I.fdesc.upload(desc, s);
    I.fpiv.upload(piv, s);
    I.fpbase.upload(pbase, s);
    I.plan.alloc(std::max<long long>(pbase[n], 1));
    k_ilu_plan<<<n, 64, 0, s>>>(I.fdesc.p, n, I.cols.p, I.srt_cols.p, I.srt_pos.p, I.fpiv.p,
                                I.fpbase.p, I.plan.p);
    HIPCHECK(hipGetLastError());
  }
  // strip sweeps: a level of more rows than the staged sweep's row groups (512 threads, 4
  // lanes per row) is split, so that no level needs a second round, whose entries would be
  // loaded inside the level instead of one level ahead
  int cap = level_cap;
  if (const char* e = getenv("RMX_LEVEL_CAP")) cap = std::max(0, atoi(e));
  std::vector<int> olo, oup;
  build_sweep(I.lo, I, true, olo, s, cap);
  build_sweep(I.up, I, false, oup, s, cap);
  build_blocks(I, group, s);
  // forward position (global, concatenated) of the row at each backward position
  std::vector<int> fpos(n), f2b(n);
  for (int r = 0; r < n; r++) fpos[olo[r]] = r;
  for (int r = 0; r < n; r++) f2b[r] = fpos[oup[r]];
  I.fwd2bwd.upload(f2b, s);
  HIPCHECK(hipStreamSynchronize(s));
  I.ready = true;
}

// ---------------------------------------------------------------------------------------
// factor / apply

template <class VT> static DBuf<VT>& LUbuf(Ilu& I);
template <> DBuf<double>& LUbuf<double>(Ilu& I) { return I.LU64; }
template <> DBuf<float>& LUbuf<float>(Ilu& I) { return I.LU32; }
template <class VT> static DBuf<VT>& PLbuf(Ilu& I);
template <> DBuf<double>& PLbuf<double>(Ilu& I) { return I.plo64; }
template <> DBuf<float>& PLbuf<float>(Ilu& I) { return I.plo32; }
template <class VT> static DBuf<VT>& PUbuf(Ilu& I);
template <> DBuf<double>& PUbuf<double>(Ilu& I) { return I.pup64; }
template <> DBuf<float>& PUbuf<float>(Ilu& I) { return I.pup32; }
// block sweep values of one precision: entries (lower, upper), unit couplings (lower, upper)
template <class VT> struct BSvals {
  DBuf<VT> *vlo, *vup, *wlo, *wup;
  DBuf<double>* rdg;
};
template <class VT> static BSvals<VT> bsvals(BlockSweep& B);
template <> BSvals<double> bsvals<double>(BlockSweep& B)
{
  return { &B.vlo64, &B.vup64, &B.wlo64, &B.wup64, &B.rdg64 };
}
template <> BSvals<float> bsvals<float>(BlockSweep& B)
{
  return { &B.vlo32, &B.vup32, &B.wlo32, &B.wup32, &B.rdg32 };
}
template <class VT> static DBuf<VT>& PDbuf(Ilu& I);
template <> DBuf<double>& PDbuf<double>(Ilu& I) { return I.pdg64; }
template <> DBuf<float>& PDbuf<float>(Ilu& I) { return I.pdg32; }

template <class VT>
static void factor_t(Ctx& c, Ilu& I, const VT* A)
{
  DBuf<VT>& LU = LUbuf<VT>(I);
  LU.alloc(I.nnz);
  HIPCHECK(hipMemcpyAsync(LU.p, A, sizeof(VT) * I.nnz, hipMemcpyDeviceToDevice, c.stream));
  int* err = reinterpret_cast<int*>(c.red.p + 8);
  HIPCHECK(hipMemsetAsync(err, 0, sizeof(int), c.stream));
  const long long plan_n = (long long)I.plan.n;
  const char* lv = getenv("RMX_ILU_LEVELS");
  if (!(lv && atoi(lv) == 1)) {
    // one persistent launch: rows in level order, pivot rows awaited by completion flags
    if (I.done.n != (size_t)I.n) {
      I.done.alloc(I.n);
      HIPCHECK(hipMemsetAsync(I.done.p, 0, sizeof(int) * I.n, c.stream));
      I.epoch = 0;
    }
    I.flow_state.alloc(4);
    if (++I.epoch == INT_MAX) {                   // epochs exhausted: clear the flags
      HIPCHECK(hipMemsetAsync(I.done.p, 0, sizeof(int) * I.n, c.stream));
      I.epoch = 1;
    }
    // one-wave workgroups, three per CU: enough rows in flight that a row is taken before
    // its pivot rows finish (about three levels ahead on the 2M-triangle meander), while
    // more waves lengthen every row (FP32 factorisation there: 640 waves 178 ms, 768
    // 168 ms, 896 182 ms, 1024 245 ms)
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, c.device));
    int per_cu = 0;
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_ilu_flow<VT>,
                                                          64, 0));
    per_cu = std::max(1, std::min(per_cu, 3));
    int nwg = (int)std::min<long long>(I.n, (long long)per_cu * prop.multiProcessorCount);
    if (const char* e = getenv("RMX_ILU_WAVES")) nwg = std::max(1, std::min(nwg, atoi(e)));
    int spin = 1 << 22;
    if (const char* e = getenv("RMX_ILU_SPIN")) spin = std::max(32, atoi(e));
    // a pass factors every row after skip; a pivot below 1e-20 is repaired as the
    // reference does (k_ilu_repair, on the first such row: every row before it is final),
    // and the rows after it are factored again from A
    int skip = -1;
    I.repaired = 0;
    for (;;) {
      int st[4] = {0, 0, 0, INT_MAX};
      HIPCHECK(hipMemcpyAsync(I.flow_state.p, st, sizeof(int) * 4, hipMemcpyHostToDevice,
                              c.stream));
      const int kt = kt_begin(c, KT_FACTOR, 3.0 * (double)I.nnz * (sizeof(VT) + 4));
      k_ilu_flow<VT><<<nwg, 64, 0, c.stream>>>(I.fdesc.p, I.n, I.cols.p, I.srt_cols.p,
                                               I.srt_pos.p, I.fpiv.p, I.fpbase.p, I.plan.p, LU.p,
                                               I.done.p, I.epoch, I.flow_state.p, spin, skip);
      kt_end(c, kt);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(st, I.flow_state.p, sizeof(int) * 4, hipMemcpyDeviceToHost,
                              c.stream));
      HIPCHECK(hipStreamSynchronize(c.stream));
      if (st[FLOW_ABORT])
        fail(RMX_ERR_HIP, "ILU(0): a row's wait for its pivot rows exceeded its bound (dataflow "
                          "factorisation aborted)");
      if (!st[FLOW_ERR]) break;
      const int bad = st[FLOW_BADROW];
      if (bad <= skip || bad >= I.n) fail(RMX_ERR_HIP, "ILU(0): inconsistent pivot record");
      k_ilu_repair<VT><<<1, 64, 0, c.stream>>>(bad, I.rp.p, I.cols.p, I.srt_cols.p, I.srt_pos.p,
                                               A, LU.p, I.flow_state.p);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(st, I.flow_state.p, sizeof(int) * 4, hipMemcpyDeviceToHost,
                              c.stream));
      HIPCHECK(hipStreamSynchronize(c.stream));
      if (st[FLOW_ERR] == 2)
        fail(RMX_ERR_PIVOT, "ILU(0) cannot be applied to the matrix: pivot of row " +
                            std::to_string(bad) + " below 1e-20 and no later row of its pattern "
                            "holds its column (PRECO_ILU0::Factor - 1)");
      I.repaired++;
      // the rows after it: the matrix again, not finished
      const long long q1 = I.h_rp[bad + 1];
      if (I.nnz > q1)
        HIPCHECK(hipMemcpyAsync(LU.p + q1, A + q1, sizeof(VT) * (I.nnz - q1),
                                hipMemcpyDeviceToDevice, c.stream));
      if (bad + 1 < I.n)
        k_flags_clear<<<(I.n - bad - 1 + 255) / 256, 256, 0, c.stream>>>(bad + 1, I.n, I.done.p);
      HIPCHECK(hipGetLastError());
      skip = bad;
    }
  } else {
  const int kt = kt_begin(c, KT_FACTOR, 3.0 * (double)I.nnz * (sizeof(VT) + 4));
  for (size_t l = 0; l + 1 < I.flev.size(); l++) {
    const int a = I.flev[l], cnt = I.flev[l + 1] - a;
    const bool last = l + 2 >= I.flev.size();
    const int an = last ? a : I.flev[l + 1];
    const int cntn = last ? 0 : I.flev[l + 2] - an;
    k_ilu_level<VT><<<cnt, 64, 0, c.stream>>>(I.fdesc.p + a, cnt, I.cols.p, I.srt_cols.p,
                                              I.srt_pos.p, I.fpiv.p, I.fpbase.p + a,
                                              I.plan.p, LU.p, err,
                                              I.fdesc.p + an, cntn, I.fpbase.p + an, plan_n);
  }
  kt_end(c, kt);
  }
  HIPCHECK(hipGetLastError());
  DBuf<VT>& pl = PLbuf<VT>(I);
  DBuf<VT>& pu = PUbuf<VT>(I);
  DBuf<VT>& pd = PDbuf<VT>(I);
  pl.alloc(std::max<long long>(I.lo.ne, 1));
  pu.alloc(std::max<long long>(I.up.ne, 1));
  pd.alloc(I.n);
  if (I.lo.ne) k_pack<VT><<<4096, 256, 0, c.stream>>>(I.lo.ne, I.lo.psrc.p, LU.p, pl.p);
  if (I.up.ne) k_pack<VT><<<4096, 256, 0, c.stream>>>(I.up.ne, I.up.psrc.p, LU.p, pu.p);
  k_pack_diag<VT><<<(I.n + 255) / 256, 256, 0, c.stream>>>(I.n, I.up.prow.p, I.rp.p, LU.p, pd.p);
  if (I.bs.on) {                                // block sweeps: values in their entry order
    BlockSweep& B = I.bs;
    const BSvals<VT> V = bsvals<VT>(B);
    const std::pair<DBuf<VT>*, const DBuf<int>*> packs[4] = {
      { V.vlo, &B.slo }, { V.vup, &B.sup }, { V.wlo, &B.sulo }, { V.wup, &B.suup } };
    for (const auto& pk : packs) {
      pk.first->alloc(pk.second->n);
      k_pack<VT><<<4096, 256, 0, c.stream>>>((long long)pk.second->n, pk.second->p, LU.p,
                                             pk.first->p);
    }
    V.rdg->alloc(B.sdg.n);
    k_pack_rcp<VT><<<4096, 256, 0, c.stream>>>((long long)B.sdg.n, B.sdg.p, LU.p, V.rdg->p);
  }
  HIPCHECK(hipGetLastError());
  int h_err = 0;
  HIPCHECK(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  if (h_err)
    fail(RMX_ERR_PIVOT, "ILU(0): pivot below 1e-20 (the per-level factorisation, "
                        "RMX_ILU_LEVELS=1, does not repair pivots; the dataflow one does)");
}

void ilu_factor(Ctx& c, Ilu& I, int prec)
{
  if (!I.ready) fail(RMX_ERR_ARG, "ILU(0): no schedule");
  if (prec == RMX_F32) {
    if (!c.have32) fail(RMX_ERR_ARG, "ILU(0) FP32: no FP32 matrix (assemble or convert first)");
    factor_t<float>(c, I, c.A32.p);
    I.fac32 = true;
  } else {
    if (!c.have64) fail(RMX_ERR_ARG, "ILU(0) FP64: no FP64 matrix");
    factor_t<double>(c, I, c.A64.p);
    I.fac64 = true;
  }
}

template <class VT, int V>
static void launch_sweeps(Ctx& c, Ilu& I, const double* b, double* x, hipEvent_t* ev,
                          const int* aidx, const double* aval)
{
  // V = 1: staged pipeline, 4 lanes per row with 6 register entries each, 512 threads (two
  // workgroups per compute unit by registers and LDS); V = 0: one-deep prefetch, 8 lanes
  // per row, 1024 threads
  const int nb = I.nblk;
  const int nt = STAGE_THREADS;
  if (ev) HIPCHECK(hipEventRecord(ev[0], c.stream));
  const double vb = (double)sizeof(VT) + 4.0;
  int kt = kt_begin(c, KT_SWEEP_LO, (double)I.lo.ne * vb + 20.0 * I.n);
  if constexpr (V == 1)
    k_sweep_stage<VT, true, 4, 12><<<nb, nt, 0, c.stream>>>(I.lo.blev.p, I.lo.bpos.p,
        I.lo.lev_ptr.p, I.lo.prs.p, I.lo.pcol.p, PLbuf<VT>(I).p, nullptr, I.lo.prow.p, b, nullptr,
        nullptr, I.lo.xp.p, nullptr, nullptr);
  else
    k_sweep<VT, true, SG, SKR><<<nb, 1024, 0, c.stream>>>(I.lo.blev.p, I.lo.bpos.p,
        I.lo.lev_ptr.p, I.lo.prs.p, I.lo.pcol.p, PLbuf<VT>(I).p, nullptr, I.lo.prow.p, b, nullptr,
        nullptr, I.lo.xp.p);
  kt_end(c, kt);
  if (ev) HIPCHECK(hipEventRecord(ev[1], c.stream));
  kt = kt_begin(c, KT_SWEEP_UP, (double)I.up.ne * vb + 24.0 * I.n);
  if constexpr (V == 1)
    k_sweep_stage<VT, false, 4, 6><<<nb, nt, 0, c.stream>>>(I.up.blev.p, I.up.bpos.p,
        I.up.lev_ptr.p, I.up.prs.p, I.up.pcol.p, PUbuf<VT>(I).p, PDbuf<VT>(I).p, I.fwd2bwd.p,
        I.lo.xp.p, I.up.prow.p, x, I.up.xp.p, aidx, aval);
  else
    k_sweep<VT, false, SG, SKR><<<nb, 1024, 0, c.stream>>>(I.up.blev.p, I.up.bpos.p,
        I.up.lev_ptr.p, I.up.prs.p, I.up.pcol.p, PUbuf<VT>(I).p, PDbuf<VT>(I).p, I.fwd2bwd.p,
        I.lo.xp.p, I.up.prow.p, x, I.up.xp.p);
  kt_end(c, kt);
  if (ev) HIPCHECK(hipEventRecord(ev[2], c.stream));
  HIPCHECK(hipGetLastError());
}

// Both sweeps of every block in one launch (one wave per block).
template <class VT>
static void launch_blocks(Ctx& c, Ilu& I, const double* b, double* x, hipEvent_t* ev,
                          const int* aidx, const double* aval)
{
  BlockSweep& B = I.bs;
  const BSvals<VT> V = bsvals<VT>(B);
  if (ev) HIPCHECK(hipEventRecord(ev[0], c.stream));
  // algorithmic bytes: entries (value + 2-byte column), per node its two records, 8 coupling
  // values and 3 FP64 reciprocal diagonals (stored as 4), per row the gather index,
  // right-hand side and result, per level its table entry, and the folded prolongation
  const double v = (double)sizeof(VT);
  const double bytes = (double)(B.ne_lo + B.ne_up) * (v + 2.0) + (double)B.nunit * (32.0 + 8.0 * v) +
                       20.0 * I.n + 4.0 * B.nlev + (aidx ? 12.0 * I.n : 0.0);
  const int kt = kt_begin(c, KT_SWEEP_BLK, bytes);
  k_block_sweep<VT><<<B.nb, 64, B.lds, c.stream>>>(B.info.p, B.gidx.p, B.lev.p,
      V.vlo->p, B.mlo.p, B.ulo.p, V.wlo->p, V.vup->p, B.mup.p, B.uup.p, V.wup->p,
      reinterpret_cast<const double4*>(V.rdg->p), b, x, aidx, aval);
  kt_end(c, kt);
  if (ev) {
    HIPCHECK(hipEventRecord(ev[1], c.stream));
    HIPCHECK(hipEventRecord(ev[2], c.stream));
  }
  HIPCHECK(hipGetLastError());
}

// RMX_SWEEP selects the sweep kernels (tuning and comparison): 0 the one-deep strip sweep,
// 1 the staged strip sweep, 2 (default) the block sweeps where the layout exists, else 1.
static int sweep_variant()
{
  const char* e = getenv("RMX_SWEEP");
  return e ? atoi(e) : 2;
}

template <class VT>
static bool apply_t(Ctx& c, Ilu& I, const double* b, double* x, hipEvent_t* ev, const int* aidx,
                    const double* aval)
{
  const int v = sweep_variant();
  if (v >= 2 && I.bs.on) { launch_blocks<VT>(c, I, b, x, ev, aidx, aval); return aidx != nullptr; }
  if (v != 0) { launch_sweeps<VT, 1>(c, I, b, x, ev, aidx, aval); return aidx != nullptr; }
  launch_sweeps<VT, 0>(c, I, b, x, ev, nullptr, nullptr);
  return false;
}

bool ilu_apply(Ctx& c, Ilu& I, int prec, const double* b, double* x, hipEvent_t* ev,
               const int* aidx, const double* aval)
{
  if (prec == RMX_F32) {
    if (!I.fac32) fail(RMX_ERR_ARG, "ILU(0) FP32 not factored");
    return apply_t<float>(c, I, b, x, ev, aidx, aval);
  }
  if (!I.fac64) fail(RMX_ERR_ARG, "ILU(0) FP64 not factored");
  return apply_t<double>(c, I, b, x, ev, aidx, aval);
}

// ---- the context's two instances ----

// Segments per strip: the requested count, or blocks of about 560 rows (at most 32 per
// strip).  Blocks much smaller than that drop most of the factor's couplings from the
// sweeps: on a 23k-equation meander reach, 32 segments of its 11 strips (66-row blocks)
// left FGMRES stalled where blocks of ~700 rows converge.
static int strip_segments(const Ctx& c, int nblk)
{
  if (c.segments > 0) return c.segments;
  const long long rows = c.n / std::max(1, nblk);
  return (int)std::max<long long>(1, std::min<long long>(32, rows / 560));
}

void setup_ilu_schedule(Ctx& c)
{
  std::vector<int> hrp, hcols;
  if (c.csr_only) {
    hrp.resize(c.n + 1); hcols.resize(c.nnz);
    c.rowptr.download(hrp.data(), c.n + 1, c.stream);
    c.cols.download(hcols.data(), c.nnz, c.stream);
    HIPCHECK(hipStreamSynchronize(c.stream));
  }
  const std::vector<int>& rp = c.csr_only ? hrp : c.hm.rowptr;
  const std::vector<int>& cl = c.csr_only ? hcols : c.hm.cols;
  ilu_setup(c.ilu, c.n, rp, cl, 1, nullptr, c.stream, nullptr, 1, 0);
  // strip ILU: two strips per compute unit (two sweep workgroups fit one CU's LDS), at
  // least a few thousand rows each
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, c.device));
  int nblk = c.strips > 0 ? c.strips : 2 * prop.multiProcessorCount;
  nblk = std::max(1, std::min(nblk, c.n / 2048));
  std::vector<int> group;
  if (!c.csr_only) {
    group.assign(c.n, -1);
    for (size_t k = 0; k < c.hm.eqno.size(); k++)
      if (c.hm.eqno[k] >= 0) group[c.hm.eqno[k]] = (int)(k / 3);
  }
  ilu_setup(c.silu, c.n, rp, cl, nblk, c.csr_only ? nullptr : &group, c.stream,
            c.csr_only ? nullptr : &c.hm, strip_segments(c, nblk), STAGE_ROWS);
  c.coarse.ready = c.coarse.fac = false;
  const int cseg = coarse_segments_of(c);
  if (!c.csr_only && cseg > 0) coarse_setup(c, rp, cl, group, cseg);
}

Ilu& ilu_of(Ctx& c, int kind)
{
  return (kind == RMX_PREC_SILU0 || kind == RMX_PREC_SILU0C || kind == RMX_PREC_MPNN) ? c.silu : c.ilu;
}

void dev_ilu0_factor(Ctx& c, int prec) { ilu_factor(c, c.ilu, prec); }

void dev_ilu0_apply(Ctx& c, int prec, const double* b, double* x)
{
  ilu_apply(c, c.ilu, prec, b, x, nullptr);
}

// Average kernel durations (HIP events on the context stream) of SpMV and of the two
// sweeps of an ILU instance, `reps` times each on the current system.
void dev_profile(Ctx& c, int prec, int reps, rmx_kernel_times& out)
{
  hipEvent_t ev[3];
  for (auto& e : ev) HIPCHECK(hipEventCreate(&e));
  double* xv = c.w[0].p;
  double* yv = c.w[1].p;
  double t_spmv = 0, t_fwd = 0, t_bwd = 0, t_sf = 0, t_sb = 0;
  dev_spmv(c, prec, c.B.p, yv);         // untimed: brings a re-laid-out matrix up to date
  for (int r = 0; r < reps; r++) {
    HIPCHECK(hipEventRecord(ev[0], c.stream));
    dev_spmv(c, prec, c.B.p, yv);
    HIPCHECK(hipEventRecord(ev[1], c.stream));
    HIPCHECK(hipEventSynchronize(ev[1]));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    t_spmv += ms * 1e-3;
    for (int k = 0; k < 2; k++) {
      Ilu& I = k == 0 ? c.ilu : c.silu;
      if (!((prec == RMX_F32) ? I.fac32 : I.fac64)) continue;
      ilu_apply(c, I, prec, c.B.p, xv, ev);
      HIPCHECK(hipEventSynchronize(ev[2]));
      float m1 = 0, m2 = 0;
      HIPCHECK(hipEventElapsedTime(&m1, ev[0], ev[1]));
      HIPCHECK(hipEventElapsedTime(&m2, ev[1], ev[2]));
      if (k == 0) { t_fwd += m1 * 1e-3; t_bwd += m2 * 1e-3; }
      else { t_sf += m1 * 1e-3; t_sb += m2 * 1e-3; }
    }
  }
  for (auto& e : ev) HIPCHECK(hipEventDestroy(e));
  out.spmv_s = t_spmv / reps;
  out.ilu_fwd_s = t_fwd / reps;
  out.ilu_bwd_s = t_bwd / reps;
  out.ilu_apply_s = out.ilu_fwd_s + out.ilu_bwd_s;
  out.silu_fwd_s = t_sf / reps;
  out.silu_bwd_s = t_sb / reps;
  out.n_spmv = reps;
  out.n_ilu = reps;
  out.spmv_bytes = prec == RMX_F32 ? spmv_f32_bytes(c)
                                   : (double)c.nnz * 12.0 + 4.0 * (c.n + 1) + 16.0 * c.n;
}

}  // namespace rmx
