// Drying and rewetting on the device, method 2 of the reference (GRID::DryRewet,
// DryRewet.cpp:696-900, called by MODEL::DoDryRewet, DoDryRew.cpp:50-238).
//
// The reference walks nodes, then elements, then the elements again for the water level of
// the marsh nodes, then the midside nodes.  Here every pass is a kernel over nodes or
// elements.  The one pass whose result depends on the element order (a marsh node takes the
// level of the wet neighbours in each wet element that holds it, the last element winning,
// and its bed is lowered at every element whose level lies below zor + dryLimit) runs per
// node over the node's elements in ascending order, which is the reference's order for that
// node.  It reads the water level of wet neighbours only: in a wet element at least one
// corner is not dry, hence not marsh, and the neighbour read when both neighbours are marsh
// (the opposite corner of a quadrilateral) is that corner, so no value read is written in
// the same pass and the result equals the sequential one bit for bit.
#include "rmx_internal.h"

namespace rmx {

// Per node: the dry flag at the start of the call kept (WD_WASDRY), the marsh flag cleared,
// dry below dryLimit (the count down restarts), else wet once the count down has run out;
// the bed back at its original level.
__global__ void k_dw_nodes(int np, const double* __restrict__ S, const double* __restrict__ zor,
                           double* __restrict__ z, uint8_t* __restrict__ nwd,
                           int* __restrict__ cdown, double dryLimit, int countDown)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  uint8_t f = nwd[n];
  f = (uint8_t)((f & ~(WD_WASDRY | WD_MARSH)) | ((f & WD_DRY) ? WD_WASDRY : 0));
  if (S[n] - zor[n] < dryLimit) {
    f |= WD_DRY;
    cdown[n] = countDown;
  } else if (cdown[n] > 0) {
    cdown[n]--;
  } else {
    f &= (uint8_t)~WD_DRY;
  }
  z[n] = zor[n];
  nwd[n] = f;
}

// Per region element: dry when all corners are dry, marsh when some are.
__global__ void k_dw_elems(int ne, const int* __restrict__ elem, const int* __restrict__ nnd,
                           const uint8_t* __restrict__ nwd, uint8_t* __restrict__ ewd)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int ncn = nnd[e] == 6 ? 3 : 4;
  int ndry = 0;
  for (int i = 0; i < ncn; i++) ndry += (nwd[elem[8LL * e + i]] & WD_DRY) ? 1 : 0;
  uint8_t f = (ewd[e] & EL_DRY) ? EL_WASDRY : 0;
  if (ndry == ncn) f |= EL_DRY;
  else if (ndry) f |= EL_MARSH;
  ewd[e] = f;
}

// position of node n in element e (-1 if absent)
__device__ __forceinline__ int pos_in(const int* nd, int k, int n)
{
  for (int i = 0; i < k; i++)
    if (nd[i] == n) return i;
  return -1;
}

// Per node: marsh if dry and a corner of a marsh element; then dry unless a wet element
// holds it.
__global__ void k_dw_flags(int np, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                           const int* __restrict__ elem, const int* __restrict__ nnd,
                           const uint8_t* __restrict__ ewd, uint8_t* __restrict__ nwd)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  uint8_t f = nwd[n];
  bool marsh = false, wet = false;
  for (int q = n2e_ptr[n]; q < n2e_ptr[n + 1]; q++) {
    const int e = n2e[q];
    const uint8_t ef = ewd[e];
    if (!(ef & EL_DRY)) wet = true;
    if ((f & WD_DRY) && (ef & EL_MARSH)) {
      const int k = nnd[e], i = pos_in(elem + 8LL * e, k, n);
      if (i >= 0 && i < (k == 6 ? 3 : 4)) marsh = true;
    }
  }
  f = (uint8_t)((f & ~(WD_DRY | WD_MARSH)) | (wet ? 0 : WD_DRY) | (marsh ? WD_MARSH : 0));
  nwd[n] = f;
}

// Per marsh corner node, over its wet elements in ascending order: a node that was dry at
// the call's start (or at the first call) takes the mean level of its two neighbouring
// corners in the element, or the one of them that is not marsh, capped at zor + rewetLimit;
// the bed is lowered to S - dryLimit where the level lies below zor + dryLimit.
__global__ void k_dw_marsh(int np, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                           const int* __restrict__ elem, const int* __restrict__ nnd,
                           const uint8_t* __restrict__ ewd, const uint8_t* __restrict__ nwd,
                           const double* __restrict__ zor, double* __restrict__ S,
                           double* __restrict__ z, double dryLimit, double rewetLimit, int first)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  const uint8_t f = nwd[n];
  if (!(f & WD_MARSH)) return;
  const bool reset = (f & WD_WASDRY) || first;
  const double zo = zor[n];
  double s = S[n], zz = z[n];
  for (int q = n2e_ptr[n]; q < n2e_ptr[n + 1]; q++) {
    const int e = n2e[q];
    if (ewd[e] & EL_DRY) continue;
    const int* nd = elem + 8LL * e;
    const int ncn = nnd[e] == 6 ? 3 : 4;
    const int i = pos_in(nd, ncn, n);
    if (i < 0) continue;
    if (reset) {
      const int j = nd[(i + 1) % ncn], k = nd[(i + ncn - 1) % ncn];
      const bool mj = nwd[j] & WD_MARSH, mk = nwd[k] & WD_MARSH;
      double si;
      if (!mj && !mk) si = (S[j] + S[k]) / 2.0;
      else if (!mj) si = S[j];
      else if (!mk) si = S[k];
      else si = S[nd[(i + 2) % ncn]];
      s = si;
      if (s > zo + rewetLimit) s = zo + rewetLimit;
      if (s < zo + dryLimit) zz = s - dryLimit;
    } else if (s < zo + dryLimit) {
      zz = s - dryLimit;
    }
  }
  S[n] = s;
  z[n] = zz;
}

// Per midside node, over all its elements in ascending order: level and bed the mean of
// the edge's corners; marsh where a marsh element's interpolated bed lies more than
// dryLimit / 10 below the original one.
__global__ void k_dw_mids(int np, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                          const int* __restrict__ elem, const int* __restrict__ nnd,
                          const uint8_t* __restrict__ ewd, uint8_t* __restrict__ nwd,
                          const double* __restrict__ zor, double* __restrict__ S,
                          double* __restrict__ z, double dryLimit)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  bool mid = false, marsh = false;
  double s = 0.0, zz = 0.0;
  for (int q = n2e_ptr[n]; q < n2e_ptr[n + 1]; q++) {
    const int e = n2e[q];
    const int* nd = elem + 8LL * e;
    const int k = nnd[e], ncn = k == 6 ? 3 : 4;
    const int i = pos_in(nd, k, n);
    if (i < ncn) continue;
    const int l = nd[i - ncn], r = nd[(i - ncn + 1) % ncn];
    s = 0.5 * (S[l] + S[r]);
    zz = 0.5 * (z[l] + z[r]);
    mid = true;
    if ((ewd[e] & EL_MARSH) && zz + dryLimit / 10.0 < zor[n]) marsh = true;
  }
  if (!mid) return;
  S[n] = s;
  z[n] = zz;
  if (marsh) nwd[n] |= WD_MARSH;
}

// Dry nodes: at rest, level and bed at the original bed.
__global__ void k_dw_reset(int np, const uint8_t* __restrict__ nwd, const double* __restrict__ zor,
                           double* __restrict__ U, double* __restrict__ V, double* __restrict__ S,
                           double* __restrict__ z, double* __restrict__ dUdt,
                           double* __restrict__ dVdt, double* __restrict__ dSdt)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np || !(nwd[n] & WD_DRY)) return;
  U[n] = 0.0; V[n] = 0.0;
  S[n] = zor[n]; z[n] = zor[n];
  dUdt[n] = 0.0; dVdt[n] = 0.0; dSdt[n] = 0.0;
}

// The end of a time step (Compute.cpp:773-783): the previous dry and marsh flags.
__global__ void k_dw_commit(int np, uint8_t* __restrict__ nwd)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  const uint8_t f = nwd[n];
  nwd[n] = (uint8_t)((f & ~(WD_DRYPREV | WD_MARSHPREV)) | ((f & WD_DRY) ? WD_DRYPREV : 0) |
                     ((f & WD_MARSH) ? WD_MARSHPREV : 0));
}

void dev_dry_rewet(Ctx& c)
{
  const int np = c.hm.np, ne = c.hm.ne;
  const int gn = (np + 255) / 256, ge = (ne + 255) / 256;
  const double dl = c.par.dry_limit;
  const double rl = c.par.rewet_limit > dl ? c.par.rewet_limit : dl;   // not below dryLimit
  hipStream_t s = c.stream;
  k_dw_nodes<<<gn, 256, 0, s>>>(np, c.S.p, c.zor.p, c.z.p, c.d_nwd.p, c.d_cdown.p, dl,
                                c.par.count_down);
  k_dw_elems<<<ge, 256, 0, s>>>(ne, c.d_elem.p, c.d_nnd.p, c.d_nwd.p, c.d_ewd.p);
  k_dw_flags<<<gn, 256, 0, s>>>(np, c.n2e_ptr.p, c.n2e.p, c.d_elem.p, c.d_nnd.p, c.d_ewd.p,
                                c.d_nwd.p);
  k_dw_marsh<<<gn, 256, 0, s>>>(np, c.n2e_ptr.p, c.n2e.p, c.d_elem.p, c.d_nnd.p, c.d_ewd.p,
                                c.d_nwd.p, c.zor.p, c.S.p, c.z.p, dl, rl, c.first_dry ? 1 : 0);
  k_dw_mids<<<gn, 256, 0, s>>>(np, c.n2e_ptr.p, c.n2e.p, c.d_elem.p, c.d_nnd.p, c.d_ewd.p,
                               c.d_nwd.p, c.zor.p, c.S.p, c.z.p, dl);
  k_dw_reset<<<gn, 256, 0, s>>>(np, c.d_nwd.p, c.zor.p, c.U.p, c.V.p, c.S.p, c.z.p, c.dUdt.p,
                                c.dVdt.p, c.dSdt.p);
  HIPCHECK(hipGetLastError());
  c.d_nwd.download(c.hm.nwd.data(), np, s);
  c.d_ewd.download(c.hm.ewd.data(), ne, s);
  HIPCHECK(hipStreamSynchronize(s));
  c.first_dry = false;
}

void dev_wetdry_commit(Ctx& c)
{
  const int np = c.hm.np;
  k_dw_commit<<<(np + 255) / 256, 256, 0, c.stream>>>(np, c.d_nwd.p);
  HIPCHECK(hipGetLastError());
}

}  // namespace rmx
